// ============================================================================
// afmm_math.h -- per-point expansion math shared by the gfx950 kernels and
// the host-side table builder / CPU model tests (all __host__ __device__).
//
// Convention: the reference's Greengard-Rokhlin normalisation, no
// Condon-Shortley phase (FmmExpansions3D.hpp:12-38):
//   Y_n^m = sqrt((n-|m|)!/(n+|m|)!) P_n^|m|(cos t) e^{i m phi}, Y_n^{-m} = conj
//   regular  R_n^m(v) = |v|^n Y_n^m,   singular S_n^m(v) = Y_n^m / |v|^{n+1}
//   P2M  M_n^m = sum q R_n^{-m}(d)           (FmmExpansions3D.hpp:148-163)
//   P2L  L_n^m = sum q S_n^{-m}(d)           (AdaptiveFMM3D.hpp:117-128)
//   potential sum M S(x-c)  /  sum L R(x-c)  (:166-177, :203-213)
//
// Storage: real sources give M_n^{-m} = conj(M_n^m) (same for L), so an
// expansion is (p+1)^2 reals, degree-major: [Re a_n^0, Re a_n^1, Im a_n^1, ...].
// Scaling: a box of half-width h stores Mhat_n^m = M_n^m / h^n and
// Lhat_n^m = L_n^m h^{n+1}; point vectors enter divided by h. Every
// translation operator is then independent of the level.
//
// Harmonics come from the Cartesian recurrences of the seminormalised
// Legendre functions (no trigonometry): with r2 = |v|^2,
//   R_0^0 = 1, R_a^a = sqrt((2a-1)/(2a)) (x+iy) R_{a-1}^{a-1},
//   R_{a+1}^a = sqrt(2a+1) z R_a^a,
//   R_n^a = ((2n-1) z R_{n-1}^a - sqrt((n-1)^2-a^2) r2 R_{n-2}^a)/sqrt(n^2-a^2)
// and the same for S with z -> z/r2, r2 -> 1/r2, x+iy -> (x+iy)/r2, S_0^0 = 1/|v|.
// They are streamed (a outer, n inner) so only three complex values are live.
// ============================================================================
#pragma once

#include <hip/hip_runtime.h>

#define AFMM_HD __host__ __device__ __forceinline__

namespace afmm {

constexpr int kMaxOrder = 16;

// Real-storage offset of Re a_n^a (Im a_n^a follows it for a > 0).
AFMM_HD constexpr int rid(int n, int a) { return n * n + (a == 0 ? 0 : 2 * a - 1); }
AFMM_HD constexpr int nterms(int p) { return (p + 1) * (p + 1); }

// ---- streamed solid harmonics ----------------------------------------------
// f(n, a, re, im) is called for every 0 <= a <= n <= L.
template <int L, class F>
AFMM_HD void regular_stream(double x, double y, double z, F&& f) {
    const double r2 = x * x + y * y + z * z;
    double dre = 1.0, dim = 0.0;  // R_a^a
#pragma unroll
    for (int a = 0; a <= L; ++a) {
        if (a > 0) {
            const double c = __builtin_sqrt((2.0 * a - 1.0) / (2.0 * a));
            const double tre = c * (x * dre - y * dim);
            const double tim = c * (x * dim + y * dre);
            dre = tre;
            dim = tim;
        }
        f(a, a, dre, dim);
        if (a + 1 <= L) {
            const double c1 = __builtin_sqrt(2.0 * a + 1.0) * z;
            double p2re = dre, p2im = dim;
            double p1re = c1 * dre, p1im = c1 * dim;
            f(a + 1, a, p1re, p1im);
#pragma unroll
            for (int n = a + 2; n <= L; ++n) {
                const double inv = 1.0 / __builtin_sqrt(double(n * n - a * a));
                const double ca = (2.0 * n - 1.0) * inv;
                const double cb = __builtin_sqrt(double((n - 1) * (n - 1) - a * a)) * inv;
                const double zz = ca * z, rr = cb * r2;
                const double qre = zz * p1re - rr * p2re;
                const double qim = zz * p1im - rr * p2im;
                f(n, a, qre, qim);
                p2re = p1re, p2im = p1im;
                p1re = qre, p1im = qim;
            }
        }
    }
}

// f(n, a, re, im) for every 0 <= a <= n <= L of S_n^a(v); v must be nonzero.
template <int L, class F>
AFMM_HD void singular_stream(double x, double y, double z, F&& f) {
    const double r2 = x * x + y * y + z * z;
    const double ir2 = 1.0 / r2;
    const double ir = __builtin_sqrt(ir2);
    const double ux = x * ir2, uy = y * ir2, uz = z * ir2;
    double dre = ir, dim = 0.0;  // S_a^a
#pragma unroll
    for (int a = 0; a <= L; ++a) {
        if (a > 0) {
            const double c = __builtin_sqrt((2.0 * a - 1.0) / (2.0 * a));
            const double tre = c * (ux * dre - uy * dim);
            const double tim = c * (ux * dim + uy * dre);
            dre = tre;
            dim = tim;
        }
        f(a, a, dre, dim);
        if (a + 1 <= L) {
            const double c1 = __builtin_sqrt(2.0 * a + 1.0) * uz;
            double p2re = dre, p2im = dim;
            double p1re = c1 * dre, p1im = c1 * dim;
            f(a + 1, a, p1re, p1im);
#pragma unroll
            for (int n = a + 2; n <= L; ++n) {
                const double inv = 1.0 / __builtin_sqrt(double(n * n - a * a));
                const double ca = (2.0 * n - 1.0) * inv;
                const double cb = __builtin_sqrt(double((n - 1) * (n - 1) - a * a)) * inv;
                const double zz = ca * uz, rr = cb * ir2;
                const double qre = zz * p1re - rr * p2re;
                const double qim = zz * p1im - rr * p2im;
                f(n, a, qre, qim);
                p2re = p1re, p2im = p1im;
                p1re = qre, p1im = qim;
            }
        }
    }
}

// ---- per-point operators ----------------------------------------------------
// P2M contribution of one body (mass q, scaled offset d from the box centre):
// Mhat_n^a += q conj(R_n^a(d)).
template <int P, class Acc>
AFMM_HD void p2m_point(double q, double dx, double dy, double dz, Acc&& add) {
    regular_stream<P>(dx, dy, dz, [&](int n, int a, double re, double im) {
        add(rid(n, a), q * re);
        if (a) add(rid(n, a) + 1, -q * im);
    });
}

// P2L contribution of one source body: Lhat_n^a += q conj(S_n^a(d)).
template <int P, class Acc>
AFMM_HD void p2l_point(double q, double dx, double dy, double dz, Acc&& add) {
    singular_stream<P>(dx, dy, dz, [&](int n, int a, double re, double im) {
        add(rid(n, a), q * re);
        if (a) add(rid(n, a) + 1, -q * im);
    });
}

// Gradient of sum_{n<=P} L_n^m R_n^m at the scaled point d. This is the
// reference's localField (FmmExpansions3D.hpp:215-235) folded onto m >= 0:
// with G+ = sum L_n^m c+_R(n,m) R_{n-1}^{m+1} the minus-branch sum equals
// conj(G+), so grad = (Re G+, Im G+, sum L c_z R).
template <int P, class Coef>
AFMM_HD void local_grad(Coef L, double dx, double dy, double dz, double& gx, double& gy,
                        double& gz) {
    double pr = 0.0, pi = 0.0, zz = 0.0;
    if constexpr (P >= 1) {
        regular_stream<P - 1>(dx, dy, dz, [&](int np, int a, double re, double im) {
            const int n = np + 1;
            {
                const double lr = L(rid(n, a)), li = a ? L(rid(n, a) + 1) : 0.0;
                const double w = (a ? 2.0 : 1.0) * __builtin_sqrt(double(n * n - a * a));
                zz += w * (re * lr - im * li);
            }
            if (a >= 1) {
                const double lr = L(rid(n, a - 1)), li = (a - 1) ? L(rid(n, a - 1) + 1) : 0.0;
                const double c = -__builtin_sqrt(double((n - a + 1) * (n - a)));
                pr += c * (re * lr - im * li);
                pi += c * (re * li + im * lr);
            }
            {
                const double lr = L(rid(n, a + 1)), li = L(rid(n, a + 1) + 1);
                const double c = __builtin_sqrt(double((n + a + 1) * (n + a)));
                pr += c * (re * lr - im * li);
                pi -= c * (re * li + im * lr);
            }
        });
    }
    gx = pr;
    gy = pi;
    gz = zz;
}

// Gradient of sum_{n<=P} M_n^m S_n^m at the scaled point d: the reference's
// multipoleField (FmmExpansions3D.hpp:180-200) folded onto m >= 0 the same way.
// With POT, *pot also receives the potential sum_{n<=P} M_n^m S_n^m
// (multipolePotential, :166-177) from the same harmonic stream.
template <int P, bool POT = false, class Coef>
AFMM_HD void multipole_grad(Coef M, double dx, double dy, double dz, double& gx, double& gy,
                            double& gz, double* pot = nullptr) {
    double pr = 0.0, pi = 0.0, zz = 0.0, ps = 0.0;
    singular_stream<P + 1>(dx, dy, dz, [&](int n2, int a, double re, double im) {
        if constexpr (POT) {
            if (n2 <= P) {
                const double mr = M(rid(n2, a)), mi = a ? M(rid(n2, a) + 1) : 0.0;
                ps += (a ? 2.0 : 1.0) * (re * mr - im * mi);
            }
        }
        if (n2 == 0) return;
        const int n = n2 - 1;
        if (a <= n) {
            const double mr = M(rid(n, a)), mi = a ? M(rid(n, a) + 1) : 0.0;
            const double w = -(a ? 2.0 : 1.0) * __builtin_sqrt(double((n + 1) * (n + 1) - a * a));
            zz += w * (re * mr - im * mi);
        }
        if (a >= 1) {
            const double mr = M(rid(n, a - 1)), mi = (a - 1) ? M(rid(n, a - 1) + 1) : 0.0;
            const double c = -__builtin_sqrt(double((n + a) * (n + a + 1)));
            pr += c * (re * mr - im * mi);
            pi += c * (re * mi + im * mr);
        }
        if (a + 1 <= n) {
            const double mr = M(rid(n, a + 1)), mi = M(rid(n, a + 1) + 1);
            const double c = __builtin_sqrt(double((n - a) * (n - a + 1)));
            pr += c * (re * mr - im * mi);
            pi -= c * (re * mi + im * mr);
        }
    });
    gx = pr;
    gy = pi;
    gz = zz;
    if constexpr (POT) *pot = ps;
}

// Potentials of the expansions at the scaled point d: sum_{n<=P} L_n^m R_n^m
// (FmmExpansions3D.hpp:203-213) and sum_{n<=P} M_n^m S_n^m (:166-177), both
// real; folded onto m >= 0 as sum_n (L_n^0 R_n^0 + 2 sum_{a>=1} Re L_n^a R_n^a).
template <int P, class Coef>
AFMM_HD double local_pot(Coef L, double dx, double dy, double dz) {
    double s = 0.0;
    regular_stream<P>(dx, dy, dz, [&](int n, int a, double re, double im) {
        const double lr = L(rid(n, a)), li = a ? L(rid(n, a) + 1) : 0.0;
        s += (a ? 2.0 : 1.0) * (re * lr - im * li);
    });
    return s;
}
template <int P, class Coef>
AFMM_HD double multipole_pot(Coef M, double dx, double dy, double dz) {
    double s = 0.0;
    singular_stream<P>(dx, dy, dz, [&](int n, int a, double re, double im) {
        const double mr = M(rid(n, a)), mi = a ? M(rid(n, a) + 1) : 0.0;
        s += (a ? 2.0 : 1.0) * (re * mr - im * mi);
    });
    return s;
}

// ---- exact box geometry -----------------------------------------------------
// A box at level l with integer coordinate i (0 <= i < 2^l) along an axis has
// the exact centre c_root + h0 * (2i + 1 - 2^l) / 2^l. It is kept as an
// unevaluated double pair (hi + lo) so that body offsets x - c are accurate to
// an ulp of the offset, independent of |c|.
struct Centre {
    double hi, lo;
};

AFMM_HD Centre exact_centre(double croot, double h0, int level, int i) {
    const double g = double(2 * (long long)i + 1 - (1LL << level));
    const double ph = h0 * g;
    const double pl = __builtin_fma(h0, g, -ph);  // h0*g = ph + pl exactly
    const double s = __builtin_ldexp(1.0, -level);
    const double a = croot, b = ph * s;
    const double sh = a + b;
    const double bb = sh - a;
    const double sl = (a - (sh - bb)) + (b - bb) + pl * s;
    const double hi = sh + sl;
    return Centre{hi, sl - (hi - sh)};
}

AFMM_HD double offset_from(double x, Centre c) { return (x - c.hi) - c.lo; }

// Offset index of an M2L pair: target minus source in units of the box edge,
// each component in [-3, 3].
AFMM_HD constexpr int offset_index(int dx, int dy, int dz) {
    return (dx + 3) * 49 + (dy + 3) * 7 + (dz + 3);
}
constexpr int kNumOffsets = 343;

// The order in which an M2L workgroup steps through the offsets ("slots"). A
// tile's targets are boxes of one octant at one level, so they lie on a
// lattice of spacing 2, and their sources at offset o lie on that lattice
// shifted by o: offsets of one parity class (o mod 2) reach the same source
// lattice, offsets of different classes disjoint ones. Slots run class by
// class, and inside a class along a boustrophedon path in which consecutive
// offsets differ by 2 in one component, so consecutive offsets' sources are
// mostly the same boxes shifted by one lattice step (kept in the L2 between
// them) instead of a disjoint set at every step.
struct M2LSlots {
    int offset_of[kNumOffsets];  // slot -> offset_index
    int slot_of[kNumOffsets];    // offset_index -> slot
};
constexpr M2LSlots make_m2l_slots() {
    M2LSlots t{};
    int k = 0;
    for (int cls = 0; cls < 8; ++cls) {
        const int par[3] = {cls >> 2, (cls >> 1) & 1, cls & 1};
        int vals[3][4] = {}, nv[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a)
            for (int d = -3; d <= 3; ++d)
                if (((d + 4) & 1) == par[a]) vals[a][nv[a]++] = d;
        int line = 0;
        for (int ix = 0; ix < nv[0]; ++ix) {
            for (int j = 0; j < nv[1]; ++j) {
                const int iy = (ix & 1) ? nv[1] - 1 - j : j;
                for (int l = 0; l < nv[2]; ++l) {
                    const int iz = (line & 1) ? nv[2] - 1 - l : l;
                    const int o = offset_index(vals[0][ix], vals[1][iy], vals[2][iz]);
                    t.offset_of[k] = o;
                    t.slot_of[o] = k;
                    ++k;
                }
                ++line;
            }
        }
    }
    return t;
}
constexpr M2LSlots kM2LSlotsHost = make_m2l_slots();

// M2L operator tiling for v_mfma_f64_16x16x4_f64: RB row blocks of 16 and the
// k steps of 4 split into NKC chunks of KC steps, one chunk of all row blocks
// (RB*KC*64 doubles, RB*KC <= 60) small enough to double-buffer in under
// 64 KB of LDS.
AFMM_HD constexpr int m2l_rb(int S) { return (S + 15) / 16; }
AFMM_HD constexpr int m2l_kb(int S) { return (S + 3) / 4; }
AFMM_HD constexpr int m2l_kcmax(int S) { return 60 / m2l_rb(S) > 0 ? 60 / m2l_rb(S) : 1; }
AFMM_HD constexpr int m2l_nkc(int S) { return (m2l_kb(S) + m2l_kcmax(S) - 1) / m2l_kcmax(S); }
AFMM_HD constexpr int m2l_kc(int S) { return (m2l_kb(S) + m2l_nkc(S) - 1) / m2l_nkc(S); }
AFMM_HD constexpr int m2l_chunk(int S) { return m2l_rb(S) * m2l_kc(S) * 64; }

// M2L operator tiling (FP64 and split-FP32 MFMA). When S leaves 1-4 rows past
// a multiple of 16 (p = 5, 6, 8, 9, ...), those tail rows run on the VALU
// instead of a 16-row MFMA block that would be mostly padding: MRB row blocks
// on MFMA plus TAIL rows. The k steps of 4 split into NKC chunks of KC steps
// (MRB*KC <= 60, the split with the fewest padded steps); a chunk is MRB*KC
// A fragments of 64 values followed by the tail rows' TAIL*KC*4 values.
AFMM_HD constexpr int m2l_mrb(int S) { return (S > 16 && S % 16 != 0 && S % 16 <= 4) ? S / 16 : (S + 15) / 16; }
AFMM_HD constexpr int m2l_tail(int S) { return S > 16 * m2l_mrb(S) ? S - 16 * m2l_mrb(S) : 0; }
AFMM_HD constexpr int m2l_mnkc(int S) {
    const int kb = m2l_kb(S), kcmax = 60 / m2l_mrb(S) > 0 ? 60 / m2l_mrb(S) : 1;
    const int n0 = (kb + kcmax - 1) / kcmax;
    int best = n0;
    for (int n = n0; n <= n0 + 1 && n <= kb; ++n)
        if (n * ((kb + n - 1) / n) < best * ((kb + best - 1) / best)) best = n;
    return best;
}
AFMM_HD constexpr int m2l_mkc(int S) { return (m2l_kb(S) + m2l_mnkc(S) - 1) / m2l_mnkc(S); }
AFMM_HD constexpr int m2l_mchunk(int S) { return (m2l_mrb(S) * 64 + m2l_tail(S) * 4) * m2l_mkc(S); }

// M2L operator tiling for v_mfma_f32_16x16x32_bf16 with operands split into
// three bf16 pieces: RB row blocks of 16, k blocks of 32 grouped into NKC
// chunks of KC blocks. A fragment set (row block, k block) is 3 pieces x 64
// lanes x 8 bf16 = 3072 bytes; RB*KC <= 24 keeps two chunk buffers under
// 150 KB of LDS.
constexpr int kBfPieces = 3;
AFMM_HD constexpr int m2l_bf_kb(int S) { return (S + 31) / 32; }
AFMM_HD constexpr int m2l_bf_kcmax(int S) { return 24 / m2l_rb(S) > 0 ? 24 / m2l_rb(S) : 1; }
AFMM_HD constexpr int m2l_bf_nkc(int S) { return (m2l_bf_kb(S) + m2l_bf_kcmax(S) - 1) / m2l_bf_kcmax(S); }
AFMM_HD constexpr int m2l_bf_kc(int S) { return (m2l_bf_kb(S) + m2l_bf_nkc(S) - 1) / m2l_bf_nkc(S); }
// 16-byte units (8 bf16) per chunk
AFMM_HD constexpr int m2l_bf_chunk(int S) { return m2l_rb(S) * m2l_bf_kc(S) * kBfPieces * 64; }

}  // namespace afmm
