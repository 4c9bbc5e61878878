// ============================================================================
// oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY. extern "C" surface of the CPU
// restatement (liboracle.so), loaded by tests/ through ctypes. The compiled
// reference (oracle/ref_shim.cpp -> oracle/_ref/libref.so) exports the same
// functions with the ref_ prefix, so the two can be compared call for call.
// Expansions cross the boundary as interleaved complex: 2*(p+1)^2 doubles.
// ============================================================================
#include <omp.h>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "oracle.hpp"
#include "tree_export.hpp"

using namespace oracle;

namespace {
V3 v3(const double* a) { return V3{a[0], a[1], a[2]}; }
Expansion unpack(const double* in, int p, V3 c) {
    Expansion e(c, p);
    for (std::size_t s = 0; s < e.a.size(); ++s) e.a[s] = cplx(in[2 * s], in[2 * s + 1]);
    return e;
}
void pack(const Expansion& e, double* out) {
    for (std::size_t s = 0; s < e.a.size(); ++s) {
        out[2 * s] = e.a[s].real();
        out[2 * s + 1] = e.a[s].imag();
    }
}

// ---- reference operator-convergence driver (src3d/fmm3d_tests.cpp) --------
// Geometry helpers restated from fmm3d_tests.cpp:85-108. `dirFirst` selects
// which operand of "scalar * randDir(rng)" draws from the generator first;
// the reference leaves that order to the compiler.
struct Geo {
    bool dirFirst;
    V3 randDir(std::mt19937& rng) const {  // :85-89
        std::uniform_real_distribution<double> u(0, 1);
        const double tau = 2.0 * std::acos(-1.0);
        const double ct = 2 * u(rng) - 1, st = std::sqrt(std::max(0.0, 1 - ct * ct)),
                     ph = tau * u(rng);
        return V3{st * std::cos(ph), st * std::sin(ph), ct};
    }
    V3 scaled(std::mt19937& rng, double (*radial)(double, double, double), double a, double b) const {
        std::uniform_real_distribution<double> u(0, 1);
        double s;
        V3 d;
        if (dirFirst) {
            d = randDir(rng);
            s = radial(u(rng), a, b);
        } else {
            s = radial(u(rng), a, b);
            d = randDir(rng);
        }
        return V3{s * d.x, s * d.y, s * d.z};
    }
    static double ballR(double u, double rad, double) { return rad * std::cbrt(u); }
    static double shellR(double u, double rmin, double rmax) { return rmin + (rmax - rmin) * u; }
    void ball(V3 c, double rad, int n, unsigned seed, std::vector<V3>& pos, std::vector<double>& q) const {
        std::mt19937 rng(seed);  // :90-98
        std::uniform_real_distribution<double> u(0, 1);
        pos.clear();
        q.clear();
        for (int i = 0; i < n; ++i) {
            const V3 d = scaled(rng, ballR, rad, 0.0);
            pos.push_back(V3{c.x + d.x, c.y + d.y, c.z + d.z});
            q.push_back(0.5 + u(rng));
        }
    }
    std::vector<V3> shell(V3 c, double rmin, double rmax, int n, unsigned seed) const {
        std::mt19937 rng(seed);  // :102-108
        std::vector<V3> pts;
        for (int i = 0; i < n; ++i) {
            const V3 d = scaled(rng, shellR, rmin, rmax);
            pts.push_back(V3{c.x + d.x, c.y + d.y, c.z + d.z});
        }
        return pts;
    }
};

double directPot(V3 t, const std::vector<V3>& s, const std::vector<double>& q) {  // :36-40
    double phi = 0.0;
    for (std::size_t j = 0; j < s.size(); ++j) {
        const V3 d = sub(t, s[j]);
        phi += q[j] / std::sqrt(dot3(d, d));
    }
    return phi;
}
std::vector<V3> directAcc(const std::vector<V3>& sp, const std::vector<double>& sq,
                          const std::vector<V3>& tp) {  // :42-51
    const std::size_t ns = sp.size(), nt = tp.size();
    std::vector<double> pos(3 * (ns + nt)), m(ns + nt, 0.0), acc(3 * (ns + nt));
    for (std::size_t i = 0; i < ns; ++i) {
        pos[3 * i] = sp[i].x, pos[3 * i + 1] = sp[i].y, pos[3 * i + 2] = sp[i].z;
        m[i] = sq[i];
    }
    for (std::size_t j = 0; j < nt; ++j)
        pos[3 * (ns + j)] = tp[j].x, pos[3 * (ns + j) + 1] = tp[j].y, pos[3 * (ns + j) + 2] = tp[j].z;
    direct(pos.data(), m.data(), ns + nt, 1.0, 0.0, acc.data());
    std::vector<V3> out(nt);
    for (std::size_t j = 0; j < nt; ++j)
        out[j] = V3{acc[3 * (ns + j)], acc[3 * (ns + j) + 1], acc[3 * (ns + j) + 2]};
    return out;
}
Expansion localFromSources(V3 zL, const std::vector<V3>& s, const std::vector<double>& q, int p) {
    Expansion out(zL, p);  // :55-68
    std::vector<cplx> Y;
    for (std::size_t j = 0; j < s.size(); ++j) {
        const double rho = harmonics(sub(s[j], zL), p, Y);
        double inv = 1.0 / rho;
        for (int n = 0; n <= p; ++n) {
            for (int m = -n; m <= n; ++m) out.a[sidx(n, m)] += q[j] * Y[sidx(n, -m)] * inv;
            inv /= rho;
        }
    }
    return out;
}
struct Err {
    double pot, acc;
};
Err relErr(const std::vector<double>& pT, const std::vector<double>& pA, const std::vector<V3>& aT,
           const std::vector<V3>& aA) {  // :73-82
    double pn = 0, pd = 0, an = 0, ad = 0;
    for (std::size_t i = 0; i < pT.size(); ++i) {
        pn += (pA[i] - pT[i]) * (pA[i] - pT[i]);
        pd += pT[i] * pT[i];
        const V3 d = sub(aA[i], aT[i]);
        an += dot3(d, d);
        ad += dot3(aT[i], aT[i]);
    }
    return {std::sqrt(pn / pd), std::sqrt(an / ad)};
}
template <class Pot, class Fld>
Err evalErr(Pot pot, Fld fld, const std::vector<V3>& tp, const std::vector<double>& pT,
            const std::vector<V3>& aT) {
    std::vector<double> pA;
    std::vector<V3> aA;
    for (const V3& t : tp) {
        pA.push_back(pot(t).real());
        aA.push_back(fld(t));
    }
    return relErr(pT, pA, aT, aA);
}
}  // namespace

extern "C" {

// Threads for the FMM passes (OpenMP over independent boxes; the results are
// the same bits at any count). 0 restores the count the process started with
// (OMP_NUM_THREADS, else every host core).
int orc_set_threads(int nthreads) {
    static const int initial = omp_get_max_threads();
    omp_set_num_threads(nthreads > 0 ? nthreads : initial);
    return omp_get_max_threads();
}

int orc_fmm(int64_t n, const double* pos, const double* mass, double G, double eps, int p, int cap,
            double* acc, int32_t* stats) {
    if (n < 0 || p < 1 || cap < 1) return -1;
    FmmStats st;
    adaptive_fmm(pos, mass, (std::size_t)n, G, eps, p, cap, acc, &st);
    if (stats) {
        stats[0] = st.maxLeafOcc;
        stats[1] = st.nBoxes;
        stats[2] = st.nLeaves;
        stats[3] = st.maxLevel;
    }
    return 0;
}

// VelocityVerlet3D (Integrator3D.hpp:19-40) driven by the oracle FMM:
// initialize, then nsteps of drift x += v*dt + a*(0.5*dt*dt) (per component,
// in that order), new forces, kick v += (a + a')*(0.5*dt). pos and vel are
// updated in place; acc receives the last accelerations.
int orc_leapfrog(int64_t n, double* pos, double* vel, const double* mass, double G, double eps, int p, int cap,
                 double dt, int nsteps, double* acc) {
    if (n < 0 || p < 1 || cap < 1 || nsteps < 0) return -1;
    std::vector<double> a(3 * (std::size_t)n), an(3 * (std::size_t)n);
    adaptive_fmm(pos, mass, (std::size_t)n, G, eps, p, cap, a.data(), nullptr);
    const double half_dt2 = 0.5 * dt * dt, half_dt = 0.5 * dt;
    for (int k = 0; k < nsteps; ++k) {
        for (std::size_t i = 0; i < 3 * (std::size_t)n; ++i) pos[i] = pos[i] + (vel[i] * dt + a[i] * half_dt2);
        adaptive_fmm(pos, mass, (std::size_t)n, G, eps, p, cap, an.data(), nullptr);
        for (std::size_t i = 0; i < 3 * (std::size_t)n; ++i) vel[i] = vel[i] + (a[i] + an[i]) * half_dt;
        a.swap(an);
    }
    for (std::size_t i = 0; i < 3 * (std::size_t)n; ++i) acc[i] = a[i];
    return 0;
}

int orc_direct(int64_t n, const double* pos, const double* mass, double G, double eps, double* acc) {
    if (n < 0) return -1;
    direct(pos, mass, (std::size_t)n, G, eps, acc);
    return 0;
}

// DirectSolver3D's pair formula (DirectSolver3D.hpp:25-30) for a subset of
// targets: acc[t] = G sum_{j != idx[t]} m_j d / (|d|^2 + eps^2)^{3/2}.
int orc_direct_targets(int64_t n, const double* pos, const double* mass, double G, double eps, int64_t nt,
                       const int64_t* idx, double* acc) {
    const double eps2 = eps * eps;
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t i = idx[t];
        double ax = 0, ay = 0, az = 0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            const double dx = pos[3 * j] - pos[3 * i], dy = pos[3 * j + 1] - pos[3 * i + 1],
                         dz = pos[3 * j + 2] - pos[3 * i + 2];
            const double r2 = dx * dx + dy * dy + dz * dz + eps2;
            const double s = mass[j] / (r2 * std::sqrt(r2));
            ax += s * dx, ay += s * dy, az += s * dz;
        }
        acc[3 * t] = G * ax, acc[3 * t + 1] = G * ay, acc[3 * t + 2] = G * az;
    }
    return 0;
}

// Softened potential at a subset of targets (Kernel3D.hpp:39-42 per pair):
// phi[t] = -G sum_{j != idx[t]} m_j / sqrt(|d|^2 + eps^2). Over all targets,
// 0.5 sum_i m_i phi_i is Diagnostics3D's pair-sum potential energy.
int orc_direct_potential_targets(int64_t n, const double* pos, const double* mass, double G, double eps,
                                 int64_t nt, const int64_t* idx, double* phi) {
    const double eps2 = eps * eps;
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t i = idx[t];
        double s = 0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            const double dx = pos[3 * j] - pos[3 * i], dy = pos[3 * j + 1] - pos[3 * i + 1],
                         dz = pos[3 * j + 2] - pos[3 * i + 2];
            s += mass[j] / std::sqrt(dx * dx + dy * dy + dz * dz + eps2);
        }
        phi[t] = -G * s;
    }
    return 0;
}

// computeDiagnostics (Diagnostics3D.hpp:20-38): out = {kinetic, potential,
// total, momentum xyz, angular momentum xyz}; the potential is the i<j pair
// sum in the reference's loop order.
int orc_diagnostics(int64_t n, const double* pos, const double* vel, const double* mass, double G, double eps,
                    double* out) {
    double ke = 0, pe = 0, P[3] = {0, 0, 0}, Lm[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; ++i) {
        const double* v = vel + 3 * i;
        const double* r = pos + 3 * i;
        const double m = mass[i];
        ke += 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        P[0] += m * v[0], P[1] += m * v[1], P[2] += m * v[2];
        Lm[0] += m * (r[1] * v[2] - r[2] * v[1]);
        Lm[1] += m * (r[2] * v[0] - r[0] * v[2]);
        Lm[2] += m * (r[0] * v[1] - r[1] * v[0]);
    }
    const double eps2 = eps * eps;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i + 1; j < n; ++j) {
            const double dx = pos[3 * j] - pos[3 * i], dy = pos[3 * j + 1] - pos[3 * i + 1],
                         dz = pos[3 * j + 2] - pos[3 * i + 2];
            const double r2 = dx * dx + dy * dy + dz * dz + eps2;
            pe += -G * mass[i] * mass[j] / std::sqrt(r2);
        }
    out[0] = ke, out[1] = pe, out[2] = ke + pe;
    out[3] = P[0], out[4] = P[1], out[5] = P[2];
    out[6] = Lm[0], out[7] = Lm[1], out[8] = Lm[2];
    return 0;
}

void* orc_tree_new(int64_t n, const double* pos, const double* mass, int cap, double epsFloor) {
    Tree* t = new Tree();
    t->build(pos, mass, (std::size_t)n, cap, epsFloor);
    t->buildLists();
    return t;
}
void orc_tree_free(void* h) { delete static_cast<Tree*>(h); }
void orc_tree_sizes(void* h, int64_t* out) {
    const Tree& t = *static_cast<Tree*>(h);
    oracle_export::sizes(t.box, t.px.size(), t.leaves.size(), t.coll, t.U, t.V, t.W, t.X, out);
}
void orc_tree_export(void* h, afmm_tree* out) {
    const Tree& t = *static_cast<Tree*>(h);
    oracle_export::boxes(t.box, out);
    oracle_export::bodies(t.px, t.py, t.pz, t.m, t.orig, out);
    if (out->leaves)
        for (std::size_t k = 0; k < t.leaves.size(); ++k) out->leaves[k] = t.leaves[k];
    oracle_export::csr(t.coll, out->colleague_off, out->colleague_idx);
    oracle_export::csr(t.U, out->u_off, out->u_idx);
    oracle_export::csr(t.V, out->v_off, out->v_idx);
    oracle_export::csr(t.W, out->w_off, out->w_idx);
    oracle_export::csr(t.X, out->x_off, out->x_idx);
    out->tol = t.tol;
}

double orc_harmonics(double x, double y, double z, int L, double* out) {
    std::vector<cplx> Y;
    const double r = harmonics(V3{x, y, z}, L, Y);
    for (std::size_t s = 0; s < Y.size(); ++s) {
        out[2 * s] = Y[s].real();
        out[2 * s + 1] = Y[s].imag();
    }
    return r;
}
double orc_acoef(int n, int m) { return A(n, m); }

void orc_p2m(const double* c, int ns, const double* spos, const double* q, int p, double* out) {
    std::vector<V3> sp(ns);
    std::vector<double> sq(q, q + ns);
    for (int i = 0; i < ns; ++i) sp[i] = v3(spos + 3 * i);
    pack(p2m(v3(c), sp, sq, p), out);
}
void orc_m2m(const double* in, int p, const double* oldc, const double* newc, double* out) {
    pack(m2m(unpack(in, p, v3(oldc)), v3(newc)), out);
}
void orc_m2l(const double* in, int p, const double* srcc, const double* locc, double* out) {
    pack(m2l(unpack(in, p, v3(srcc)), v3(locc)), out);
}
void orc_l2l(const double* in, int p, const double* oldc, const double* newc, double* out) {
    pack(l2l(unpack(in, p, v3(oldc)), v3(newc)), out);
}
// Potential (complex, 2 doubles) and gradient (3 doubles) of an expansion.
void orc_local_eval(const double* in, int p, const double* c, const double* t, double* pot, double* grad) {
    const Expansion e = unpack(in, p, v3(c));
    const cplx ph = local_potential(e, v3(t));
    const V3 g = local_field(e, v3(t));
    pot[0] = ph.real(), pot[1] = ph.imag();
    grad[0] = g.x, grad[1] = g.y, grad[2] = g.z;
}
void orc_multipole_eval(const double* in, int p, const double* c, const double* t, double* pot,
                        double* grad) {
    const Expansion e = unpack(in, p, v3(c));
    const cplx ph = multipole_potential(e, v3(t));
    const V3 g = multipole_field(e, v3(t));
    pot[0] = ph.real(), pot[1] = ph.imag();
    grad[0] = g.x, grad[1] = g.y, grad[2] = g.z;
}

void orc_plummer(int64_t n, double a, double M, double G, unsigned seed, double* pos, double* vel,
                 double* mass) {
    plummer((std::size_t)n, a, M, G, seed, pos, vel, mass);
}

// Operator error table of fmm3d_tests.cpp:202-288 for p = 2,4,6,8,10:
// out[(op*5 + ip)*2 + {0: potential rms, 1: acceleration rms}], op order
// P2M, M2L (worst of four geometries), M2M, L2L.
void orc_operator_convergence(int dirFirst, double* out) {
    const Geo g{dirFirst != 0};
    const int ps[5] = {2, 4, 6, 8, 10};
    auto store = [&](int op, int ip, Err e) {
        out[(op * 5 + ip) * 2] = e.pot;
        out[(op * 5 + ip) * 2 + 1] = e.acc;
    };
    std::vector<V3> sp;
    std::vector<double> sq;
    {  // P2M :202-218
        const V3 zc{0, 0, 0};
        g.ball(zc, 1.0, 80, 11, sp, sq);
        const std::vector<V3> tp = g.shell(zc, 4.0, 8.0, 64, 99);
        std::vector<double> pT;
        for (const V3& t : tp) pT.push_back(directPot(t, sp, sq));
        const std::vector<V3> aT = directAcc(sp, sq, tp);
        for (int ip = 0; ip < 5; ++ip) {
            const Expansion M = p2m(zc, sp, sq, ps[ip]);
            store(0, ip, evalErr([&](V3 t) { return multipole_potential(M, t); },
                                 [&](V3 t) { return multipole_field(M, t); }, tp, pT, aT));
        }
    }
    {  // M2L :220-250, worst over four geometries
        const V3 zLs[4] = {{6, 0, 0}, {0, 6, 0}, {0, 0, 6}, {3.5, 3.5, 3.5}};
        Err worst[5] = {};
        for (const V3& zL : zLs) {
            const V3 zc{0, 0, 0};
            g.ball(zc, 0.9, 80, 21, sp, sq);
            std::vector<V3> tp;
            std::vector<double> dummy;
            g.ball(zL, 0.9, 64, 77, tp, dummy);
            std::vector<double> pT;
            for (const V3& t : tp) pT.push_back(directPot(t, sp, sq));
            const std::vector<V3> aT = directAcc(sp, sq, tp);
            for (int ip = 0; ip < 5; ++ip) {
                const Expansion L = m2l(p2m(zc, sp, sq, ps[ip]), zL);
                const Err e = evalErr([&](V3 t) { return local_potential(L, t); },
                                      [&](V3 t) { return local_field(L, t); }, tp, pT, aT);
                worst[ip].pot = std::max(worst[ip].pot, e.pot);
                worst[ip].acc = std::max(worst[ip].acc, e.acc);
            }
        }
        for (int ip = 0; ip < 5; ++ip) store(1, ip, worst[ip]);
    }
    {  // M2M :252-269
        const V3 zChild{0.5, 0.4, 0.3}, zParent{0, 0, 0};
        g.ball(zChild, 0.4, 80, 31, sp, sq);
        const std::vector<V3> tp = g.shell(zParent, 5.0, 9.0, 64, 55);
        std::vector<double> pT;
        for (const V3& t : tp) pT.push_back(directPot(t, sp, sq));
        const std::vector<V3> aT = directAcc(sp, sq, tp);
        for (int ip = 0; ip < 5; ++ip) {
            const Expansion Mp = m2m(p2m(zChild, sp, sq, ps[ip]), zParent);
            store(2, ip, evalErr([&](V3 t) { return multipole_potential(Mp, t); },
                                 [&](V3 t) { return multipole_field(Mp, t); }, tp, pT, aT));
        }
    }
    {  // L2L :271-288
        const V3 zParent{0, 0, 0}, zChild{0.5, 0.4, 0.3};
        g.ball(V3{-8, 3, 2}, 1.0, 80, 41, sp, sq);
        std::vector<V3> tp;
        std::vector<double> dummy;
        g.ball(zChild, 0.4, 64, 66, tp, dummy);
        std::vector<double> pT;
        for (const V3& t : tp) pT.push_back(directPot(t, sp, sq));
        const std::vector<V3> aT = directAcc(sp, sq, tp);
        for (int ip = 0; ip < 5; ++ip) {
            const Expansion Lc = l2l(localFromSources(zParent, sp, sq, ps[ip]), zChild);
            store(3, ip, evalErr([&](V3 t) { return local_potential(Lc, t); },
                                 [&](V3 t) { return local_field(Lc, t); }, tp, pT, aT));
        }
    }
}

}  // extern "C"
