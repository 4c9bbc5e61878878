// ============================================================================
// afmm_tables.cpp -- host construction of the translation matrices.
//
// The complex operator entries follow the reference's Greengard-Rokhlin
// translation theorems term by term (A_n^m = (-1)^n / sqrt((n-m)!(n+m)!),
// i^e phases, FmmExpansions3D.hpp:56-67, :237-330), evaluated in long double
// from Cartesian solid-harmonic recurrences. Each complex operator T acting
// on conjugate-symmetric coefficient vectors is then written as a real S x S
// matrix on the storage of afmm_math.h and scaled to unit box size:
//   M2L: t = 2*offset (target - source), unit half-width on both sides
//   M2M: t = child - parent = s/2 (parent unit); input scaled by 2^-n
//   L2L: t = child - parent = s/2 (parent unit); output scaled by 2^-(j+1)
// ============================================================================
#include "afmm_tables.h"

#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <mutex>

#include "afmm_math.h"

namespace afmm {
namespace {

using ld = long double;
using cld = std::complex<ld>;

int cidx(int n, int m) { return n * n + n + m; }

struct Coefs {
    int L;
    std::vector<ld> A;  // A_n^m at cidx(n, m), n <= L
    explicit Coefs(int L_) : L(L_), A((L_ + 1) * (L_ + 1)) {
        for (int n = 0; n <= L; ++n)
            for (int m = -n; m <= n; ++m)
                A[cidx(n, m)] = ((n & 1) ? -1.0L : 1.0L) *
                                std::exp(-0.5L * (std::lgamma((ld)(n - m + 1)) + std::lgamma((ld)(n + m + 1))));
    }
    ld a(int n, int m) const { return A[cidx(n, m)]; }
};

ld ipow_even(int e) { return ((((e / 2) % 2) + 2) % 2 == 0) ? 1.0L : -1.0L; }

// All R_l^q(v) (regular) or S_l^q(v) (singular), l <= L, -l <= q <= l.
void solid(ld x, ld y, ld z, int L, bool singular, std::vector<cld>& out) {
    out.assign((L + 1) * (L + 1), cld(0, 0));
    const ld r2 = x * x + y * y + z * z;
    ld ux = x, uy = y, uz = z, rr = r2;
    cld diag(1, 0);
    if (singular) {
        ux = x / r2, uy = y / r2, uz = z / r2, rr = 1.0L / r2;
        diag = cld(1.0L / std::sqrt(r2), 0);
    }
    for (int a = 0; a <= L; ++a) {
        if (a > 0) diag *= std::sqrt((2.0L * a - 1) / (2.0L * a)) * cld(ux, uy);
        cld p2 = diag, p1 = std::sqrt(2.0L * a + 1) * uz * diag;
        out[cidx(a, a)] = diag;
        if (a + 1 <= L) out[cidx(a + 1, a)] = p1;
        for (int n = a + 2; n <= L; ++n) {
            const ld inv = 1.0L / std::sqrt((ld)(n * n - a * a));
            const cld q = ((2.0L * n - 1) * inv * uz) * p1 -
                          (std::sqrt((ld)((n - 1) * (n - 1) - a * a)) * inv * rr) * p2;
            out[cidx(n, a)] = q;
            p2 = p1;
            p1 = q;
        }
    }
    for (int n = 1; n <= L; ++n)
        for (int a = 1; a <= n; ++a) out[cidx(n, -a)] = std::conj(out[cidx(n, a)]);
}

// Real matrix of a complex operator entry(j,k,n,m) (output (j,k), input (n,m)),
// column-major, with row/column scale factors.
template <class Entry, class RowScale, class ColScale>
void realify(int p, Entry entry, RowScale rs, ColScale cs, double* out) {
    const int S = nterms(p);
    for (int n = 0; n <= p; ++n)
        for (int a = 0; a <= n; ++a)
            for (int part = 0; part < (a ? 2 : 1); ++part) {
                const int col = rid(n, a) + part;
                for (int j = 0; j <= p; ++j)
                    for (int k = 0; k <= j; ++k) {
                        cld v;
                        if (a == 0) v = entry(j, k, n, 0);
                        else if (part == 0) v = entry(j, k, n, a) + entry(j, k, n, -a);
                        else v = cld(0, 1) * (entry(j, k, n, a) - entry(j, k, n, -a));
                        const ld sc = rs(j) * cs(n);
                        out[(std::size_t)col * S + rid(j, k)] = (double)(sc * v.real());
                        if (k) out[(std::size_t)col * S + rid(j, k) + 1] = (double)(sc * v.imag());
                    }
            }
}

Tables build(int p) {
    Tables t;
    t.p = p;
    t.S = nterms(p);
    const int S = t.S;
    const Coefs C(2 * p + 1);
    t.m2m.assign((std::size_t)8 * S * S, 0.0);
    t.l2l.assign((std::size_t)8 * S * S, 0.0);
    t.m2l.assign((std::size_t)kNumOffsets * S * S, 0.0);
    std::vector<cld> H;
    auto one = [](int) { return 1.0L; };
    for (int o = 0; o < 8; ++o) {
        const ld sx = (o & 1) ? 0.5L : -0.5L, sy = (o & 2) ? 0.5L : -0.5L, sz = (o & 4) ? 0.5L : -0.5L;
        solid(sx, sy, sz, p, false, H);
        // M2M: M'_j^k = sum M_{n'}^{m'} i^{|k|-|k-m'|-|m'|} A_{j-n'}^{k-m'} A_{n'}^{m'}
        //               R_{j-n'}^{-(k-m')}(t) / A_j^k
        auto m2m = [&](int j, int k, int n1, int m1) -> cld {
            const int n = j - n1, m = k - m1;
            if (n < 0 || std::abs(m) > n) return cld(0, 0);
            const ld ph = ipow_even(std::abs(k) - std::abs(m) - std::abs(k - m));
            return ph * C.a(n, m) * C.a(n1, m1) * H[cidx(n, -m)] / C.a(j, k);
        };
        realify(p, m2m, one, [](int n) { return std::ldexp(1.0L, -n); },
                t.m2m.data() + (std::size_t)o * S * S);
        // L2L: L'_j^k = sum_{n>=j} L_n^m i^{|m|-|m-k|-|k|} A_{n-j}^{m-k} A_j^k
        //               R_{n-j}^{m-k}(t) / A_n^m
        auto l2l = [&](int j, int k, int n, int m) -> cld {
            const int nn = n - j, mm = m - k;
            if (nn < 0 || std::abs(mm) > nn) return cld(0, 0);
            const ld ph = ipow_even(std::abs(m) - std::abs(m - k) - std::abs(k));
            return ph * C.a(nn, mm) * C.a(j, k) * H[cidx(nn, mm)] / C.a(n, m);
        };
        realify(p, l2l, [](int j) { return std::ldexp(1.0L, -(j + 1)); }, one,
                t.l2l.data() + (std::size_t)o * S * S);
    }
    for (int dx = -3; dx <= 3; ++dx)
        for (int dy = -3; dy <= 3; ++dy)
            for (int dz = -3; dz <= 3; ++dz) {
                if (std::abs(dx) <= 1 && std::abs(dy) <= 1 && std::abs(dz) <= 1) continue;
                solid(2.0L * dx, 2.0L * dy, 2.0L * dz, 2 * p, true, H);
                // M2L: L_j^k = (-1)^j sum M_n^m i^{|k-m|-|k|-|m|} A_n^m A_j^k
                //              S_{j+n}^{m-k}(t) / A_{j+n}^{m-k}
                auto m2l = [&](int j, int k, int n, int m) -> cld {
                    const int deg = j + n, ord = m - k;
                    const ld ph = ((j & 1) ? -1.0L : 1.0L) *
                                  ipow_even(std::abs(k - m) - std::abs(k) - std::abs(m));
                    return ph * C.a(n, m) * C.a(j, k) * H[cidx(deg, ord)] / C.a(deg, ord);
                };
                realify(p, m2l, one, one,
                        t.m2l.data() + (std::size_t)offset_index(dx, dy, dz) * S * S);
            }
    return t;
}

}  // namespace

std::vector<double> mfma_a_layout(const std::vector<double>& cm, int S, int nmat) {
    const int RB = m2l_rb(S), NKC = m2l_nkc(S), KC = m2l_kc(S), CH = m2l_chunk(S);
    std::vector<double> out((size_t)nmat * NKC * CH, 0.0);
    for (int t = 0; t < nmat; ++t)
        for (int c = 0; c < NKC; ++c)
            for (int rb = 0; rb < RB; ++rb)
                for (int kk = 0; kk < KC; ++kk)
                    for (int lane = 0; lane < 64; ++lane) {
                        const int row = 16 * rb + lane % 16, col = 4 * (c * KC + kk) + lane / 16;
                        if (row < S && col < S)
                            out[((size_t)t * NKC + c) * CH + ((size_t)rb * KC + kk) * 64 + lane] =
                                cm[(size_t)t * S * S + (size_t)col * S + row];
                    }
    return out;
}

std::vector<double> m2l_a_layout(const std::vector<double>& cm, int S, int nmat) {
    const int MRB = m2l_mrb(S), TR = m2l_tail(S), NKC = m2l_mnkc(S), KC = m2l_mkc(S), CH = m2l_mchunk(S);
    std::vector<double> out((size_t)nmat * NKC * CH, 0.0);
    auto at = [&](int t, int row, int col) {
        return (row < S && col < S) ? cm[(size_t)t * S * S + (size_t)col * S + row] : 0.0;
    };
    for (int t = 0; t < nmat; ++t)
        for (int c = 0; c < NKC; ++c) {
            double* ch = out.data() + ((size_t)t * NKC + c) * CH;
            for (int rb = 0; rb < MRB; ++rb)
                for (int kk = 0; kk < KC; ++kk)
                    for (int lane = 0; lane < 64; ++lane)
                        ch[((size_t)rb * KC + kk) * 64 + lane] =
                            at(t, 16 * rb + lane % 16, 4 * (c * KC + kk) + lane / 16);
            double* tail = ch + (size_t)MRB * KC * 64;
            for (int tr = 0; tr < TR; ++tr)
                for (int kk = 0; kk < KC; ++kk)
                    for (int kq = 0; kq < 4; ++kq)
                        tail[((size_t)tr * KC + kk) * 4 + kq] = at(t, 16 * MRB + tr, 4 * (c * KC + kk) + kq);
        }
    return out;
}

namespace {
uint16_t bf16_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (uint16_t)(u >> 16);
}
double bf16_value(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return (double)f;
}
}  // namespace

void bf16_split3(double x, uint16_t out[3]) {
    double r = x;
    for (int k = 0; k < kBfPieces; ++k) {
        out[k] = bf16_bits((float)r);
        r -= bf16_value(out[k]);
    }
}

std::vector<uint16_t> mfma_bf16_layout(const std::vector<double>& cm, int S, int nmat) {
    const int RB = m2l_rb(S), KB = m2l_bf_kb(S), NKC = m2l_bf_nkc(S), KC = m2l_bf_kc(S);
    const size_t CHU = (size_t)m2l_bf_chunk(S);
    std::vector<uint16_t> out((size_t)nmat * NKC * CHU * 8, 0);
    uint16_t pc[kBfPieces];
    for (int t = 0; t < nmat; ++t)
        for (int c = 0; c < NKC; ++c)
            for (int rb = 0; rb < RB; ++rb)
                for (int kk = 0; kk < KC; ++kk) {
                    const int kb = c * KC + kk;
                    for (int lane = 0; lane < 64; ++lane)
                        for (int e = 0; e < 8; ++e) {
                            const int row = 16 * rb + lane % 16, col = 32 * kb + 8 * (lane / 16) + e;
                            const double v = (row < S && col < S && kb < KB)
                                                 ? cm[(size_t)t * S * S + (size_t)col * S + row] : 0.0;
                            bf16_split3(v, pc);
                            for (int q = 0; q < kBfPieces; ++q)
                                out[(((((size_t)t * NKC + c) * RB + rb) * KC + kk) * kBfPieces + q) * 512 +
                                    (size_t)lane * 8 + e] = pc[q];
                        }
                }
    return out;
}

const Tables& tables_for_order(int p) {
    static std::mutex mu;
    static std::map<int, Tables> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(p);
    if (it == cache.end()) it = cache.emplace(p, build(p)).first;
    return it->second;
}

}  // namespace afmm
