// ============================================================================
// oracle.cpp -- TEST INFRASTRUCTURE ONLY. CPU restatement of the reference
// adaptive FMM (see oracle.hpp for the pin list). Build: oracle/Makefile
// (g++ -O3 -std=c++17, no -march: x86-64 baseline has no FMA, so no
// contraction, exactly like the reference's own Release build).
// ============================================================================
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <random>

namespace oracle {

// ---------------------------------------------------------------------------
// Tree
// ---------------------------------------------------------------------------

// AdaptiveOctree::adjacent (AdaptiveOctree.hpp:58-62): per axis
// |dc| <= half_a + half_b + tol.
bool Tree::adjacent(int a, int b) const {
    const Box& P = box[a];
    const Box& Q = box[b];
    const double reach = P.half + Q.half + tol;
    return std::fabs(P.cx - Q.cx) <= reach && std::fabs(P.cy - Q.cy) <= reach &&
           std::fabs(P.cz - Q.cz) <= reach;
}

int Tree::maxLeafOcc() const {  // AdaptiveOctree.hpp:191-195
    int occ = 0;
    for (int b : leaves) occ = std::max(occ, box[b].pcount);
    return occ;
}

// AdaptiveOctree::build + subdivide (AdaptiveOctree.hpp:65-134), restated
// with an explicit stack. Popping children in octant order reproduces the
// recursive DFS pre-order, so box ids (children allocated as one block when
// the parent is visited) and the stable octant permutation are identical.
void Tree::build(const double* pos, const double* mass, std::size_t n, int cap, double floorEdge) {
    epsFloor = floorEdge;
    box.clear();
    leaves.clear();
    std::vector<int> perm(n);
    for (std::size_t i = 0; i < n; ++i) perm[i] = (int)i;

    // Bounding cube (:71-81). std::min/std::max semantics: replace only on a
    // strict comparison, so the first of equal values wins.
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (std::size_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            const double v = pos[3 * i + a];
            if (v < lo[a]) lo[a] = v;
            if (hi[a] < v) hi[a] = v;
        }
    const double ctr[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
    double ext = hi[0] - lo[0];
    if (ext < hi[1] - lo[1]) ext = hi[1] - lo[1];
    if (ext < hi[2] - lo[2]) ext = hi[2] - lo[2];
    double half = 0.5 * ext;
    if (!(half > 0)) half = 1.0;
    half *= 1.0 + 1e-9;
    tol = 1e-9 * half;

    Box root;
    root.cx = ctr[0];
    root.cy = ctr[1];
    root.cz = ctr[2];
    root.half = half;
    root.pcount = (int)n;
    box.push_back(root);

    std::vector<int> scratch;
    std::vector<int> todo{0};
    while (!todo.empty()) {
        const int bi = todo.back();
        todo.pop_back();
        const Box cur = box[bi];
        // Split rule (:102-104).
        if (cur.pcount <= cap || 2.0 * cur.half <= epsFloor || cur.level >= kMaxDepth) continue;
        auto octant = [&](int j) {
            return (pos[3 * j] >= cur.cx ? 1 : 0) | (pos[3 * j + 1] >= cur.cy ? 2 : 0) |
                   (pos[3 * j + 2] >= cur.cz ? 4 : 0);
        };
        // Stable counting sort of the box's index range by octant (:112-118).
        int count[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const int a0 = cur.pfirst, a1 = cur.pfirst + cur.pcount;
        for (int k = a0; k < a1; ++k) ++count[octant(perm[k])];
        int first[8], cursor[8];
        for (int o = 0, run = a0; o < 8; ++o) {
            first[o] = cursor[o] = run;
            run += count[o];
        }
        scratch.assign(perm.begin() + a0, perm.begin() + a1);
        for (int j : scratch) perm[cursor[octant(j)]++] = j;
        // Children for non-empty octants, centre = parent +- half/2 (:120-132).
        box[bi].leaf = false;
        const double h2 = 0.5 * cur.half;
        int kids[8], nk = 0;
        for (int o = 0; o < 8; ++o) {
            if (count[o] == 0) continue;
            Box ch;
            ch.parent = bi;
            ch.level = cur.level + 1;
            ch.half = h2;
            ch.cx = cur.cx + ((o & 1) ? h2 : -h2);
            ch.cy = cur.cy + ((o & 2) ? h2 : -h2);
            ch.cz = cur.cz + ((o & 4) ? h2 : -h2);
            ch.pfirst = first[o];
            ch.pcount = count[o];
            const int id = (int)box.size();
            box.push_back(ch);
            box[bi].child[o] = id;
            kids[nk++] = id;
        }
        // Recurse in octant order (:133): push reversed so the first pops first.
        for (int k = nk - 1; k >= 0; --k) todo.push_back(kids[k]);
    }

    // Gather the permuted SoA arrays (:90-95) and the leaf list (:96).
    px.resize(n);
    py.resize(n);
    pz.resize(n);
    m.resize(n);
    orig.resize(n);
    for (std::size_t k = 0; k < n; ++k) {
        const int j = perm[k];
        px[k] = pos[3 * j];
        py[k] = pos[3 * j + 1];
        pz[k] = pos[3 * j + 2];
        m[k] = mass[j];
        orig[k] = j;
    }
    for (int b = 0; b < (int)box.size(); ++b)
        if (box[b].leaf && box[b].pcount > 0) leaves.push_back(b);
}

// AdaptiveOctree::buildLists (AdaptiveOctree.hpp:137-189).
void Tree::buildLists() {
    const int nb = (int)box.size();
    coll.assign(nb, {});
    U.assign(nb, {});
    V.assign(nb, {});
    W.assign(nb, {});
    X.assign(nb, {});
    if (nb == 0) return;
    // Colleagues and V (:143-161): children of the parent's colleagues, in
    // colleague order then octant order, split by adjacency.
    coll[0].push_back(0);
    for (int b = 1; b < nb; ++b) {
        for (int pc : coll[box[b].parent])
            for (int o = 0; o < 8; ++o) {
                const int c = box[pc].child[o];
                if (c < 0) continue;
                (adjacent(b, c) ? coll[b] : V[b]).push_back(c);
            }
    }
    // U (:164, :176-180): DFS from the root through adjacent boxes, keeping
    // the non-empty leaves in visit order.
    std::vector<int> stack;
    for (int b : leaves) {
        stack.assign(1, 0);
        while (!stack.empty()) {
            const int node = stack.back();
            stack.pop_back();
            if (!adjacent(node, b)) continue;
            if (box[node].leaf) {
                if (box[node].pcount > 0) U[b].push_back(node);
                continue;
            }
            for (int o = 7; o >= 0; --o)
                if (box[node].child[o] >= 0) stack.push_back(box[node].child[o]);
        }
    }
    // W (:167-169, :181-189): descend each non-leaf colleague; a separated
    // child is emitted, an adjacent non-leaf child is descended, both in
    // octant order. The action stack holds ~id for "emit id" and id for
    // "descend into id", pushed reversed so actions run in octant order.
    for (int b : leaves)
        for (int c : coll[b]) {
            if (c == b || box[c].leaf) continue;
            stack.assign(1, c);
            while (!stack.empty()) {
                const int act = stack.back();
                stack.pop_back();
                if (act < 0) {
                    W[b].push_back(~act);
                    continue;
                }
                int next[8], na = 0;
                for (int o = 0; o < 8; ++o) {
                    const int ch = box[act].child[o];
                    if (ch < 0) continue;
                    if (!adjacent(ch, b)) next[na++] = ~ch;
                    else if (!box[ch].leaf) next[na++] = ch;
                }
                for (int k = na - 1; k >= 0; --k) stack.push_back(next[k]);
            }
        }
    // X (:172-173): transpose of W, leaves scanned in ascending id.
    for (int c : leaves)
        for (int w : W[c]) X[w].push_back(c);
}

// ---------------------------------------------------------------------------
// Harmonics and operators (FmmExpansions3D.hpp)
// ---------------------------------------------------------------------------
namespace {
struct LogFact {  // logFact (FmmExpansions3D.hpp:46-54)
    double v[256];
    LogFact() {
        v[0] = 0.0;
        for (int i = 1; i < 256; ++i) v[i] = v[i - 1] + std::log(static_cast<double>(i));
    }
};
const LogFact& logfact() {
    static const LogFact t;
    return t;
}
// Gradient coefficients (FmmExpansions3D.hpp:111-129).
double gzR(int n, int m) { return std::sqrt(double(n * n - m * m)); }
double gpR(int n, int m) { return ((m >= 0) ? -1.0 : 1.0) * std::sqrt(double((n - m) * (n - m - 1))); }
double gmR(int n, int m) { return ((m > 0) ? 1.0 : -1.0) * std::sqrt(double((n + m) * (n + m - 1))); }
double gzS(int n, int m) { return -std::sqrt(double((n + 1) * (n + 1) - m * m)); }
double gpS(int n, int m) { return ((m >= 0) ? -1.0 : 1.0) * std::sqrt(double((n + m + 1) * (n + m + 2))); }
double gmS(int n, int m) { return ((m > 0) ? 1.0 : -1.0) * std::sqrt(double((n - m + 1) * (n - m + 2))); }
}  // namespace

// Acoef (FmmExpansions3D.hpp:57-60).
double A(int n, int m) {
    const LogFact& lf = logfact();
    const double sgn = (n & 1) ? -1.0 : 1.0;
    const int a = n - m, b = n + m;
    return sgn * std::exp(-0.5 * ((a < 0 ? 0.0 : lf.v[a]) + (b < 0 ? 0.0 : lf.v[b])));
}

// jpow (FmmExpansions3D.hpp:63-67): i^e for even e.
double ipow_even(int e) {
    const long h = e / 2;
    return (((h % 2) + 2) % 2 == 0) ? 1.0 : -1.0;
}

// computeY (FmmExpansions3D.hpp:75-107): seminormalised Legendre recurrence,
// no Condon-Shortley phase; Y_n^{-m} = conj(Y_n^m). Returns |v|.
double harmonics(const V3& v, int L, std::vector<cplx>& Y) {
    const double r = std::sqrt(dot3(v, v));
    Y.assign((L + 1) * (L + 1), cplx(0, 0));
    const double ct = (r > 0) ? v.z / r : 1.0;
    const double rxy = std::sqrt(std::max(0.0, v.x * v.x + v.y * v.y));
    const double st = (r > 0) ? rxy / r : 0.0;
    const double phi = std::atan2(v.y, v.x);
    const int w = L + 1;
    std::vector<double> Q(w * w, 0.0);
    Q[0] = 1.0;
    for (int m = 1; m <= L; ++m)
        Q[m * w + m] = st * std::sqrt((2.0 * m - 1.0) / (2.0 * m)) * Q[(m - 1) * w + (m - 1)];
    for (int m = 0; m <= L; ++m) {
        if (m + 1 <= L) Q[(m + 1) * w + m] = ct * std::sqrt(2.0 * m + 1.0) * Q[m * w + m];
        for (int n = m + 2; n <= L; ++n)
            Q[n * w + m] = (ct * (2.0 * n - 1.0) * Q[(n - 1) * w + m] -
                            std::sqrt(double((n - 1) * (n - 1) - m * m)) * Q[(n - 2) * w + m]) /
                           std::sqrt(double(n * n - m * m));
    }
    for (int n = 0; n <= L; ++n)
        for (int m = 0; m <= n; ++m) {
            const cplx e{std::cos(m * phi), std::sin(m * phi)};
            const double q = Q[n * w + m];
            Y[sidx(n, m)] = q * e;
            if (m > 0) Y[sidx(n, -m)] = q * std::conj(e);
        }
    return r;
}

// P2M (FmmExpansions3D.hpp:148-163): M_n^m += q rho^n Y_n^{-m}.
Expansion p2m(V3 c, const std::vector<V3>& pos, const std::vector<double>& q, int p) {
    Expansion out(c, p);
    std::vector<cplx> Y;
    for (std::size_t i = 0; i < pos.size(); ++i) {
        const double rho = harmonics(sub(pos[i], c), p, Y);
        double rn = 1.0;
        for (int n = 0; n <= p; ++n) {
            for (int m = -n; m <= n; ++m) out.a[sidx(n, m)] += q[i] * rn * Y[sidx(n, -m)];
            rn *= rho;
        }
    }
    return out;
}

// multipolePotential (FmmExpansions3D.hpp:166-177).
cplx multipole_potential(const Expansion& mp, V3 t) {
    std::vector<cplx> Y;
    const double r = harmonics(sub(t, mp.c), mp.p, Y);
    cplx phi{0, 0};
    const double ir = 1.0 / r;
    double irn = ir;
    for (int n = 0; n <= mp.p; ++n) {
        for (int m = -n; m <= n; ++m) phi += mp.a[sidx(n, m)] * Y[sidx(n, m)] * irn;
        irn *= ir;
    }
    return phi;
}

// multipoleField (FmmExpansions3D.hpp:180-200): gradient through the
// singular-harmonic raising coefficients, degree p+1 harmonics.
V3 multipole_field(const Expansion& mp, V3 t) {
    std::vector<cplx> Y;
    const double r = harmonics(sub(t, mp.c), mp.p + 1, Y);
    std::vector<double> ir(mp.p + 3);
    ir[0] = 1.0;
    for (std::size_t k = 1; k < ir.size(); ++k) ir[k] = ir[k - 1] / r;
    auto Sg = [&](int n, int m) -> cplx {
        if (n < 0 || std::abs(m) > n) return cplx(0, 0);
        return Y[sidx(n, m)] * ir[n + 1];
    };
    cplx gx{0, 0}, gy{0, 0}, gz{0, 0};
    for (int n = 0; n <= mp.p; ++n)
        for (int m = -n; m <= n; ++m) {
            const cplx c = mp.a[sidx(n, m)];
            const cplx up = Sg(n + 1, m + 1), dn = Sg(n + 1, m - 1), mid = Sg(n + 1, m);
            gx += c * 0.5 * (gpS(n, m) * up + gmS(n, m) * dn);
            gy += c * cplx{0, -0.5} * (gpS(n, m) * up - gmS(n, m) * dn);
            gz += c * gzS(n, m) * mid;
        }
    return V3{gx.real(), gy.real(), gz.real()};
}

// localPotential (FmmExpansions3D.hpp:203-213).
cplx local_potential(const Expansion& lp, V3 t) {
    std::vector<cplx> Y;
    const double r = harmonics(sub(t, lp.c), lp.p, Y);
    cplx phi{0, 0};
    double rn = 1.0;
    for (int n = 0; n <= lp.p; ++n) {
        for (int m = -n; m <= n; ++m) phi += lp.a[sidx(n, m)] * Y[sidx(n, m)] * rn;
        rn *= r;
    }
    return phi;
}

// localField (FmmExpansions3D.hpp:215-235): gradient through the
// regular-harmonic lowering coefficients.
V3 local_field(const Expansion& lp, V3 t) {
    std::vector<cplx> Y;
    const double r = harmonics(sub(t, lp.c), lp.p, Y);
    std::vector<double> rk(lp.p + 2);
    rk[0] = 1.0;
    for (std::size_t k = 1; k < rk.size(); ++k) rk[k] = rk[k - 1] * r;
    auto Rg = [&](int n, int m) -> cplx {
        if (n < 0 || std::abs(m) > n) return cplx(0, 0);
        return Y[sidx(n, m)] * rk[n];
    };
    cplx gx{0, 0}, gy{0, 0}, gz{0, 0};
    for (int n = 0; n <= lp.p; ++n)
        for (int m = -n; m <= n; ++m) {
            const cplx c = lp.a[sidx(n, m)];
            const cplx up = Rg(n - 1, m + 1), dn = Rg(n - 1, m - 1), mid = Rg(n - 1, m);
            gx += c * 0.5 * (gpR(n, m) * up + gmR(n, m) * dn);
            gy += c * cplx{0, -0.5} * (gpR(n, m) * up - gmR(n, m) * dn);
            gz += c * gzR(n, m) * mid;
        }
    return V3{gx.real(), gy.real(), gz.real()};
}

// M2M (FmmExpansions3D.hpp:242-265), t = old centre - new centre.
Expansion m2m(const Expansion& in, V3 newc) {
    const int p = in.p;
    Expansion out(newc, p);
    std::vector<cplx> Y;
    const double rho = harmonics(sub(in.c, newc), p, Y);
    std::vector<double> rk(p + 1);
    rk[0] = 1.0;
    for (int n = 1; n <= p; ++n) rk[n] = rk[n - 1] * rho;
    for (int j = 0; j <= p; ++j)
        for (int k = -j; k <= j; ++k) {
            cplx s{0, 0};
            for (int n = 0; n <= j; ++n)
                for (int m = -n; m <= n; ++m) {
                    const int jn = j - n, km = k - m;
                    if (std::abs(km) > jn) continue;
                    const double ph = ipow_even(std::abs(k) - std::abs(m) - std::abs(k - m));
                    s += in.a[sidx(jn, km)] * ph * A(n, m) * A(jn, km) * rk[n] * Y[sidx(n, -m)];
                }
            out.a[sidx(j, k)] = s / A(j, k);
        }
    return out;
}

// M2L (FmmExpansions3D.hpp:273-299), t = local centre - source centre,
// including the (-1)^j of the local-target A_j^k (:293-296).
Expansion m2l(const Expansion& in, V3 localc) {
    const int p = in.p;
    Expansion out(localc, p);
    std::vector<cplx> Y;
    const double rho = harmonics(sub(localc, in.c), 2 * p, Y);
    std::vector<double> rk(2 * p + 2);
    rk[0] = 1.0;
    for (std::size_t n = 1; n < rk.size(); ++n) rk[n] = rk[n - 1] * rho;
    for (int j = 0; j <= p; ++j)
        for (int k = -j; k <= j; ++k) {
            cplx s{0, 0};
            for (int n = 0; n <= p; ++n)
                for (int m = -n; m <= n; ++m) {
                    const int deg = j + n, ord = m - k;
                    if (std::abs(ord) > deg) continue;
                    const double ph = ipow_even(std::abs(k - m) - std::abs(k) - std::abs(m));
                    s += in.a[sidx(n, m)] * ph * A(n, m) * A(j, k) * Y[sidx(deg, ord)] /
                         (A(deg, ord) * rk[deg + 1]);
                }
            out.a[sidx(j, k)] = ((j & 1) ? -1.0 : 1.0) * s;
        }
    return out;
}

// L2L (FmmExpansions3D.hpp:307-330), t = new centre - old centre.
Expansion l2l(const Expansion& in, V3 newc) {
    const int p = in.p;
    Expansion out(newc, p);
    std::vector<cplx> Y;
    const double rho = harmonics(sub(newc, in.c), p, Y);
    std::vector<double> rk(p + 1);
    rk[0] = 1.0;
    for (int n = 1; n <= p; ++n) rk[n] = rk[n - 1] * rho;
    for (int j = 0; j <= p; ++j)
        for (int k = -j; k <= j; ++k) {
            cplx s{0, 0};
            for (int n = j; n <= p; ++n)
                for (int m = -n; m <= n; ++m) {
                    const int nj = n - j, mk = m - k;
                    if (std::abs(mk) > nj) continue;
                    const double ph = ipow_even(std::abs(m) - std::abs(m - k) - std::abs(k));
                    s += in.a[sidx(n, m)] * ph * A(nj, mk) * A(j, k) * Y[sidx(nj, mk)] * rk[nj] /
                         A(n, m);
                }
            out.a[sidx(j, k)] = s;
        }
    return out;
}

// P2L (AdaptiveFMM3D.hpp:117-128): L_n^m += q Y_n^{-m}(r - c) / rho^{n+1}.
void p2l_accumulate(Expansion& L, const Tree& t, const Box& src, V3 cb) {
    std::vector<cplx> Y;
    for (int k = src.pfirst; k < src.pfirst + src.pcount; ++k) {
        const V3 d{t.px[k] - cb.x, t.py[k] - cb.y, t.pz[k] - cb.z};
        const double rho = harmonics(d, L.p, Y);
        double inv = 1.0 / rho;
        for (int n = 0; n <= L.p; ++n) {
            for (int m = -n; m <= n; ++m) L.a[sidx(n, m)] += t.m[k] * Y[sidx(n, -m)] * inv;
            inv /= rho;
        }
    }
}

// ---------------------------------------------------------------------------
// Solvers
// ---------------------------------------------------------------------------
void adaptive_fmm(const double* pos, const double* mass, std::size_t n, double G, double eps,
                  int p, int cap, double* acc, FmmStats* st) {
    for (std::size_t i = 0; i < 3 * n; ++i) acc[i] = 0.0;
    if (n == 0) return;  // AdaptiveFMM3D.hpp:34
    Tree t;
    t.build(pos, mass, n, cap, eps);  // eps floor (:37)
    t.buildLists();
    const int nb = (int)t.box.size();
    const int S = (p + 1) * (p + 1);
    if (st) {
        st->maxLeafOcc = t.maxLeafOcc();
        st->nBoxes = nb;
        st->nLeaves = (int)t.leaves.size();
        st->maxLevel = 0;
        for (const Box& b : t.box) st->maxLevel = std::max(st->maxLevel, b.level);
    }
    auto centre = [&](int b) { return V3{t.box[b].cx, t.box[b].cy, t.box[b].cz}; };
    std::vector<Expansion> M(nb), L(nb);
    for (int b = 0; b < nb; ++b) {
        M[b] = Expansion(centre(b), p);
        L[b] = Expansion(centre(b), p);
    }
    // Boxes by level: the reference visits them in descending (upward) and
    // ascending (downward) id, which orders every child after its parent;
    // here the boxes of one level, whose sums do not depend on each other,
    // may run on several threads (OpenMP), each box's own sums in the
    // reference's order, so the results are the same bits at any thread count.
    int maxlev = 0;
    for (const Box& bx : t.box) maxlev = std::max(maxlev, bx.level);
    std::vector<std::vector<int>> bylev(maxlev + 1);
    for (int b = 0; b < nb; ++b) bylev[t.box[b].level].push_back(b);
    const int nl = (int)t.leaves.size();
    // Upward pass (:53-71): P2M at leaves, M2M children -> parent.
#pragma omp parallel for schedule(dynamic, 16)
    for (int li = 0; li < nl; ++li) {
        const int b = t.leaves[li];
        const Box& bx = t.box[b];
        std::vector<V3> sp;
        std::vector<double> sq;
        for (int k = bx.pfirst; k < bx.pfirst + bx.pcount; ++k) {
            sp.push_back(V3{t.px[k], t.py[k], t.pz[k]});
            sq.push_back(t.m[k]);
        }
        M[b] = p2m(centre(b), sp, sq, p);
    }
    for (int lv = maxlev; lv >= 0; --lv) {
        const std::vector<int>& ids = bylev[lv];
#pragma omp parallel for schedule(dynamic, 16)
        for (int k = 0; k < (int)ids.size(); ++k) {
            const int b = ids[k];
            if (t.box[b].leaf) continue;
            for (int o = 0; o < 8; ++o) {
                const int ch = t.box[b].child[o];
                if (ch < 0) continue;
                const Expansion sh = m2m(M[ch], centre(b));
                for (int s = 0; s < S; ++s) M[b].a[s] += sh.a[s];
            }
        }
    }
    // Downward pass (:74-85): L2L from parent, M2L over V, P2L over X.
    for (int lv = 0; lv <= maxlev; ++lv) {
        const std::vector<int>& ids = bylev[lv];
#pragma omp parallel for schedule(dynamic, 16)
        for (int k = 0; k < (int)ids.size(); ++k) {
            const int b = ids[k];
            const V3 cb = centre(b);
            if (t.box[b].parent >= 0) {
                const Expansion up = l2l(L[t.box[b].parent], cb);
                for (int s = 0; s < S; ++s) L[b].a[s] += up.a[s];
            }
            for (int c : t.V[b]) {
                const Expansion tr = m2l(M[c], cb);
                for (int s = 0; s < S; ++s) L[b].a[s] += tr.a[s];
            }
            for (int c : t.X[b]) p2l_accumulate(L[b], t, t.box[c], cb);
        }
    }
    // Leaves (:88-108): L2P + M2P over W + softened P2P over U (self skipped).
    const double eps2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 8)
    for (int li = 0; li < nl; ++li) {
        const int b = t.leaves[li];
        const Box& bx = t.box[b];
        for (int ti = bx.pfirst; ti < bx.pfirst + bx.pcount; ++ti) {
            const V3 zi{t.px[ti], t.py[ti], t.pz[ti]};
            V3 far = local_field(L[b], zi);
            for (int c : t.W[b]) {
                const V3 g = multipole_field(M[c], zi);
                far.x += g.x;
                far.y += g.y;
                far.z += g.z;
            }
            V3 near{};
            for (int uc : t.U[b]) {
                const Box& cc = t.box[uc];
                for (int j = cc.pfirst; j < cc.pfirst + cc.pcount; ++j) {
                    if (j == ti) continue;
                    const V3 d{t.px[j] - zi.x, t.py[j] - zi.y, t.pz[j] - zi.z};
                    const double r2 = dot3(d, d) + eps2;
                    const double s = t.m[j] / (r2 * std::sqrt(r2));
                    near.x += s * d.x;
                    near.y += s * d.y;
                    near.z += s * d.z;
                }
            }
            const int o = t.orig[ti];
            acc[3 * o] = G * (far.x + near.x);
            acc[3 * o + 1] = G * (far.y + near.y);
            acc[3 * o + 2] = G * (far.z + near.z);
        }
    }
}

void direct(const double* pos, const double* mass, std::size_t n, double G, double eps, double* acc) {
    for (std::size_t i = 0; i < 3 * n; ++i) acc[i] = 0.0;
    const double eps2 = eps * eps;
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = i + 1; j < n; ++j) {
            const V3 d{pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1],
                       pos[3 * j + 2] - pos[3 * i + 2]};
            const double r2 = dot3(d, d) + eps2;
            const double inv = 1.0 / (r2 * std::sqrt(r2));
            const double gi = G * inv;
            const V3 base{gi * d.x, gi * d.y, gi * d.z};
            acc[3 * i] += mass[j] * base.x;
            acc[3 * i + 1] += mass[j] * base.y;
            acc[3 * i + 2] += mass[j] * base.z;
            acc[3 * j] -= mass[i] * base.x;
            acc[3 * j + 1] -= mass[i] * base.y;
            acc[3 * j + 2] -= mass[i] * base.z;
        }
}

// makePlummerSphere (InitialConditions3D.hpp:100-152): inverse-CDF radius
// with u >= 1e-12, Aarseth-Henon-Wielen speed rejection bounded by 0.1,
// isotropic directions; then COM position and momentum removed.
void plummer(std::size_t n, double a, double mtot, double G, unsigned seed, double* pos,
             double* vel, double* mass) {
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    const double mi = mtot / static_cast<double>(n);
    const double vscale = std::sqrt(2.0 * G * mtot);
    const double twopi = 2.0 * M_PI;
    auto direction = [&](double out[3]) {
        const double ct = 2.0 * uni(rng) - 1.0;
        const double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
        const double ph = twopi * uni(rng);
        out[0] = st * std::cos(ph);
        out[1] = st * std::sin(ph);
        out[2] = ct;
    };
    for (std::size_t i = 0; i < n; ++i) {
        double u = uni(rng);
        if (u < 1e-12) u = 1e-12;
        const double r = a / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
        double q, g;
        do {
            q = uni(rng);
            g = 0.1 * uni(rng);
        } while (g > q * q * std::pow(1.0 - q * q, 3.5));
        const double vesc = vscale * std::pow(r * r + a * a, -0.25);
        mass[i] = mi;
        double d[3];
        direction(d);
        for (int k = 0; k < 3; ++k) pos[3 * i + k] = r * d[k];
        direction(d);
        const double s = q * vesc;
        for (int k = 0; k < 3; ++k) vel[3 * i + k] = s * d[k];
    }
    double pc[3] = {0, 0, 0}, pm[3] = {0, 0, 0}, mt = 0.0;
    for (std::size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            pc[k] += mass[i] * pos[3 * i + k];
            pm[k] += mass[i] * vel[3 * i + k];
        }
        mt += mass[i];
    }
    const double inv = 1.0 / mt;
    for (std::size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            pos[3 * i + k] -= inv * pc[k];
            vel[3 * i + k] -= inv * pm[k];
        }
}

}  // namespace oracle
