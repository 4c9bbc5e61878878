// ============================================================================
// fmm.hip -- FMM passes on a built tree (AdaptiveFMM3D.hpp:44-108):
//   P2M at leaves -> M2M bottom-up -> M2L (V) -> P2L (X) -> L2L top-down
//   -> L2P + M2P (W) per body -> P2P (U) per leaf, acc = G*(far + near).
// M2L and P2L only read multipoles, so all levels run in one launch each;
// L2L then folds parents into children level by level, which gives the same
// sum as the reference's per-box order (parents first, AdaptiveFMM3D.hpp:74).
// Expansions use the scaled real basis of afmm_math.h; translations use the
// level-independent tables of afmm_tables.cpp (column-major S x S).
// ============================================================================
#include <algorithm>

#include "afmm_internal.h"
#include "afmm_math.h"

namespace afmm {
namespace {

constexpr int kT = 256;
inline int blocks(long long n, int t = kT) { return (int)std::max<long long>(1, (n + t - 1) / t); }

// ---- wave helpers -------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;  // every lane holds the total (fixed butterfly order)
}

// ---- P2M and P2L: per-body coefficient accumulation -------------------------------
// One wave per target box. Lanes take the source bodies 64 at a time from the
// packed stream; each lane accumulates q conj(R(d)) (P2M, FmmExpansions3D.hpp:
// 148-163) or q conj(S(d)) (P2L, AdaptiveFMM3D.hpp:117-128) of its bodies into
// a register window of the S coefficients; the windows are then summed across
// the wave in a fixed butterfly order. Orders above 8 take two windows (two
// passes over the stream) to bound registers.
//   KIND 0: targets = leaves, sources = the leaf's own bodies, Mhat = result.
//   KIND 1: targets = boxes, sources = bodies of X(box) leaves, Lhat += result.
template <int S>
constexpr int coef_window() { return S <= 96 ? S : (S + 1) / 2; }

// Bodies of a list of boxes staged through LDS in chunks of kStreamChunk list
// entries: their (first body, count) and the running body offset, so that
// stream position q maps to its body by a binary search in LDS rather than by
// walking the list in global memory.
constexpr int kStreamChunk = 256;
struct StreamLds {
    int first[kStreamChunk];
    int pre[kStreamChunk + 1];
};

// Loads entries [e0, e0 + m) of `list` (m <= kStreamChunk) and returns the
// number of bodies they hold. All 64 lanes of the (single-wave) block call it.
__device__ __forceinline__ int stage_chunk(StreamLds& sl, const int* list, int e0, int m, const int* seg_first,
                                           const int* seg_count, int lane) {
    int cnt[kStreamChunk / 64];
    int run = 0;
#pragma unroll
    for (int r = 0; r < kStreamChunk / 64; ++r) {
        const int e = lane * (kStreamChunk / 64) + r;  // lane owns 4 consecutive entries
        int c = 0;
        if (e < m) {
            const int box = list[e0 + e];
            sl.first[e] = seg_first[box];
            c = seg_count[box];
        }
        cnt[r] = run;
        run += c;
    }
    // exclusive scan of the per-lane totals
    int incl = run;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    const int base = incl - run;
#pragma unroll
    for (int r = 0; r < kStreamChunk / 64; ++r) {
        const int e = lane * (kStreamChunk / 64) + r;
        if (e < kStreamChunk) sl.pre[e] = base + cnt[r];
    }
    const int total = __shfl(incl, 63, 64);
    if (lane == 0) sl.pre[kStreamChunk] = total;
    __syncthreads();
    return total;
}

// Body of stream position q (< total) of the staged chunk with m entries.
__device__ __forceinline__ int stream_body(const StreamLds& sl, int m, int q) {
    int lo = 0, hi = m - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sl.pre[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    return sl.first[lo] + (q - sl.pre[lo]);
}

template <int P, int KIND, int C0, int CW>
__device__ __forceinline__ void accumulate_window(const TreeView& t, const int* list, int nlist, const double* ch,
                                                  double ih, int lane, StreamLds& sl, double* __restrict__ out,
                                                  size_t out_base) {
    constexpr int S = nterms(P);
    double acc[CW];
#pragma unroll
    for (int k = 0; k < CW; ++k) acc[k] = 0.0;
    for (int e0 = 0; e0 < nlist; e0 += kStreamChunk) {
        const int m = min(kStreamChunk, nlist - e0);
        __syncthreads();  // previous chunk's readers are done
        const int total = stage_chunk(sl, list, e0, m, t.seg_first, t.seg_count, lane);
        for (int q0 = 0; q0 < total; q0 += 64) {
            const int q = q0 + lane;
            if (q < total) {
                const int i = stream_body(sl, m, q);
                const double dx = offset_from(t.x[i], Centre{ch[0], ch[1]}) * ih;
                const double dy = offset_from(t.y[i], Centre{ch[2], ch[3]}) * ih;
                const double dz = offset_from(t.z[i], Centre{ch[4], ch[5]}) * ih;
                auto add = [&](int k, double v) {
                    if (k >= C0 && k < C0 + CW) acc[k - C0] += v;
                };
                if (KIND == 0) p2m_point<P>(t.m[i], dx, dy, dz, add);
                else p2l_point<P>(t.m[i], dx, dy, dz, add);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < CW; ++k) {
        const double v = wave_sum(acc[k]);
        if (lane == (k & 63) && C0 + k < S) {
            if (KIND == 0) out[out_base + C0 + k] = v;
            else out[out_base + C0 + k] += v;
        }
    }
}

template <int P, int KIND>
__global__ void __launch_bounds__(64) k_accumulate(TreeView t, double* __restrict__ out) {
    constexpr int S = nterms(P);
    constexpr int CW = coef_window<S>();
    __shared__ StreamLds sl;
    const int lane = threadIdx.x;
    int b, nlist;
    const int* list;
    if (KIND == 0) {
        b = t.leaf_ids[blockIdx.x];
        list = t.leaf_ids + blockIdx.x;
        nlist = 1;
    } else {
        b = blockIdx.x;
        const int xa = t.x_off[b];
        list = t.x_idx + xa;
        nlist = t.x_off[b + 1] - xa;
        if (nlist == 0) return;
    }
    const double ch[6] = {t.cdd[6 * (size_t)b], t.cdd[6 * (size_t)b + 1], t.cdd[6 * (size_t)b + 2],
                          t.cdd[6 * (size_t)b + 3], t.cdd[6 * (size_t)b + 4], t.cdd[6 * (size_t)b + 5]};
    const double ih = t.invh[b];
    accumulate_window<P, KIND, 0, CW>(t, list, nlist, ch, ih, lane, sl, out, (size_t)b * S);
    if constexpr (CW < S) accumulate_window<P, KIND, CW, S - CW>(t, list, nlist, ch, ih, lane, sl, out, (size_t)b * S);
}

// ---- M2P over W (AdaptiveFMM3D.hpp:95, FmmExpansions3D.hpp:180-200) --------------
// One wave per target leaf. The wave is split into G = floor(64/T) lane groups
// of T = max(n_leaf, 8) lanes (T = 64 above 64 bodies); group g handles W
// entries g, g+G, ... with their multipoles staged in LDS. Partial fields are
// reduced across groups in a fixed order and added to far[] (and, with POT,
// the potential to fpot[]).
template <int P, bool POT>
__global__ void __launch_bounds__(64) k_m2p(TreeView t, const double* __restrict__ M, double* __restrict__ far,
                                           double* __restrict__ fpot) {
    constexpr int S = nterms(P);
    __shared__ double mw[8][S];
    __shared__ double red[4][64];
    const int li = blockIdx.x;
    const int wa = t.w_off[li], nw = t.w_off[li + 1] - wa;
    if (nw == 0) return;
    const int b = t.leaf_ids[li];
    const int f = t.seg_first[b], nt = t.seg_count[b];
    const int lane = threadIdx.x;
    const int T = nt > 64 ? 64 : (nt < 8 ? 8 : nt);
    const int Gs = 64 / T;
    const int g = lane / T, tl = lane % T;
    for (int tb = 0; tb < nt; tb += T) {
        const int i = f + tb + tl;
        const bool act = g < Gs && tb + tl < nt;
        const double xi = act ? t.x[i] : 0.0, yi = act ? t.y[i] : 0.0, zi = act ? t.z[i] : 0.0;
        double fx = 0.0, fy = 0.0, fz = 0.0, fp = 0.0;
        for (int r = 0; r < nw; r += Gs) {
            __syncthreads();
            const int nq = min(Gs, nw - r);
            for (int e = lane; e < nq * S; e += 64) {
                const int q = e / S, c = e - q * S;
                mw[q][c] = M[(size_t)t.w_idx[wa + r + q] * S + c];
            }
            __syncthreads();
            if (act && g < nq) {
                const int w = t.w_idx[wa + r + g];
                const double* c = t.cdd + 6 * (size_t)w;
                const double ih = t.invh[w];
                const double dx = offset_from(xi, Centre{c[0], c[1]}) * ih;
                const double dy = offset_from(yi, Centre{c[2], c[3]}) * ih;
                const double dz = offset_from(zi, Centre{c[4], c[5]}) * ih;
                double gx, gy, gz;
                const double* mg = mw[g];
                multipole_grad<P>([&](int k) { return mg[k]; }, dx, dy, dz, gx, gy, gz);
                const double s2 = ih * ih;
                fx += gx * s2;
                fy += gy * s2;
                fz += gz * s2;
                if constexpr (POT) fp += multipole_pot<P>([&](int k) { return mg[k]; }, dx, dy, dz) * ih;
            }
        }
        if (Gs > 1) {
            __syncthreads();
            red[0][lane] = fx;
            red[1][lane] = fy;
            red[2][lane] = fz;
            red[3][lane] = fp;
            __syncthreads();
            if (g == 0)
                for (int q = 1; q < Gs; ++q) {
                    fx += red[0][q * T + tl];
                    fy += red[1][q * T + tl];
                    fz += red[2][q * T + tl];
                    fp += red[3][q * T + tl];
                }
        }
        if (act && g == 0) {
            far[3 * (size_t)i] += fx;
            far[3 * (size_t)i + 1] += fy;
            far[3 * (size_t)i + 2] += fz;
            if constexpr (POT) fpot[i] += fp;
        }
    }
}

// ---- M2L on FP64 matrix cores ------------------------------------------------
// L_t = sum_{c in V(t)} T[offset(t,c)] M_c with level-independent S x S
// operators. One wave owns 16 targets (the 16 columns of
// v_mfma_f64_16x16x4_f64) and loops over the 316 separated offsets; for each
// offset present in its tile it streams the pre-swizzled operator (A, one
// contiguous 512-byte fragment per MFMA) against the 16 gathered source
// multipoles (B, zero where a target has no source at that offset), keeping
// the 16 x S output tile in accumulators for the whole loop. Targets are
// ordered by (level, octant) so that a tile shares most of its offset pattern.
// Operand layouts (CDNA 16x16x4 f64): A[i][k] in lane i + 16k, B[k][j] in lane
// j + 16k, D[i][j] in lane j + 16(i%4), register i/4 (rows interleave by 4,
// unlike the f32 16x16x4 form; see ck's mfma_type for f64 16x16x4:
// group_size 1, num_groups_per_blk 4).
typedef double dbl4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(kT) k_rank_of(int nb, const int* __restrict__ order, int* __restrict__ rank) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i < nb) rank[order[i]] = i;
}

__global__ void __launch_bounds__(kT) k_fill_int(int* a, size_t n, int v) {
    const size_t i = blockIdx.x * (size_t)kT + threadIdx.x;
    if (i < n) a[i] = v;
}

// srcT[offset * np + rank(t)] = source box of target t at that offset.
__global__ void __launch_bounds__(kT) k_m2l_src(int nb, int np, const int* __restrict__ rank, const int* __restrict__ vl,
                                               const int* __restrict__ nv, const int* __restrict__ ic,
                                               int* __restrict__ srcT) {
    const int b = blockIdx.x * kT + threadIdx.x;
    if (b >= nb) return;
    const int r = rank[b];
    for (int e = 0; e < nv[b]; ++e) {
        const int c = vl[kVSlots * (size_t)b + e];
        const int o = offset_index(ic[3 * (size_t)b] - ic[3 * (size_t)c], ic[3 * (size_t)b + 1] - ic[3 * (size_t)c + 1],
                                   ic[3 * (size_t)b + 2] - ic[3 * (size_t)c + 2]);
        srcT[(size_t)o * np + r] = c;
    }
}

constexpr int kM2LWaves = 8;
template <int P>
constexpr int m2l_cb() { return P <= 8 ? 2 : 1; }  // 16-target column blocks per wave
template <int P>
constexpr int m2l_tile() { return kM2LWaves * 16 * m2l_cb<P>(); }

// present[wg][o] = 1 when some target of workgroup tile wg has a source at
// offset o (one wave per (tile, offset); the tile's ranks are contiguous in srcT).
__global__ void __launch_bounds__(64) k_m2l_plan(int tile, int np, int ntgt, const int* __restrict__ srcT,
                                                unsigned char* __restrict__ present) {
    const int wg = blockIdx.x / kNumOffsets, o = blockIdx.x % kNumOffsets;
    const int dx = o / 49 - 3, dy = (o / 7) % 7 - 3, dz = o % 7 - 3;
    const bool near = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
    bool any = false;
    if (!near)
        for (int k = threadIdx.x; k < tile; k += 64) {
            const int r = wg * tile + k;
            if (r < ntgt && srcT[(size_t)o * np + r] >= 0) any = true;
        }
    const bool f = __ballot(any) != 0ull;
    if (threadIdx.x == 0) present[(size_t)wg * kNumOffsets + o] = f ? 1 : 0;
}

typedef float flt4 __attribute__((ext_vector_type(4)));

// Operator fragment and accumulator types. FP64: v_mfma_f64_16x16x4f64 on
// double operators. Split FP32: each operand x is carried as hi = fl32(x),
// lo = fl32(x - hi), and a 16x16x4 product is three v_mfma_f32_16x16x4f32
// (lo*hi + hi*lo + hi*hi) into an FP32 accumulator.
template <bool MIX>
struct M2LTypes {
    using A = double;
    using Acc = dbl4;
};
template <>
struct M2LTypes<true> {
    using A = float2;
    using Acc = flt4;
};

template <int P, bool MIX>
__global__ void __launch_bounds__(kM2LWaves * 64) k_m2l_mfma(int nb, int np, int ntgt,
                                                           const int* __restrict__ order,
                                                           const int* __restrict__ srcT,
                                                           const unsigned char* __restrict__ present_g,
                                                           const typename M2LTypes<MIX>::A* __restrict__ Asw,
                                                           const double* __restrict__ M, double* __restrict__ L) {
    using AT = typename M2LTypes<MIX>::A;
    using Acc = typename M2LTypes<MIX>::Acc;
    constexpr int S = nterms(P);
    constexpr int RB = m2l_rb(S), NKC = m2l_nkc(S), KC = m2l_kc(S), CH = m2l_chunk(S);
    constexpr int CB = m2l_cb<P>();
    constexpr int TILE = m2l_tile<P>();
    constexpr int NT = kM2LWaves * 64;
    constexpr int PER = (CH + NT - 1) / NT;
    __shared__ AT abuf[2][CH];
    __shared__ unsigned char present[kNumOffsets + 1];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int j = lane & 15, kq = lane >> 4;
    const int base = blockIdx.x * TILE;
    int rank[CB];
    bool valid[CB];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
        rank[cb] = base + (wave * CB + cb) * 16 + j;
        valid[cb] = rank[cb] < ntgt;
    }
    Acc acc[CB][RB];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[cb][rb][r] = 0;
    if (base < ntgt) {
        for (int t = tid; t < kNumOffsets; t += NT) present[t] = present_g[(size_t)blockIdx.x * kNumOffsets + t];
        if (tid == 0) present[kNumOffsets] = 1;  // sentinel
        __syncthreads();
        auto next_present = [&](int o) {
            ++o;
            while (!present[o]) ++o;
            return o;  // kNumOffsets when none left
        };
        // Software pipeline over (offset o, k chunk c) items: while the MFMAs of
        // the current item run, the next item's operator chunk (global -> regs
        // -> other LDS buffer) and its gathered source multipoles are in flight.
        auto load_b = [&](const int (&src)[CB], int c, double (&bv)[CB][KC]) {
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) {
                const double* ms = M + (size_t)(src[cb] >= 0 ? src[cb] : 0) * S;
#pragma unroll
                for (int kk = 0; kk < KC; ++kk) {
                    const int k = 4 * (c * KC + kk) + kq;
                    bv[cb][kk] = (src[cb] >= 0 && k < S) ? ms[k] : 0.0;
                }
            }
        };
        auto load_src = [&](int o, int (&src)[CB]) {
            bool any = false;
#pragma unroll
            for (int cb = 0; cb < CB; ++cb) {
                src[cb] = (o < kNumOffsets && valid[cb]) ? srcT[(size_t)o * np + rank[cb]] : -1;
                any |= src[cb] >= 0;
            }
            return __ballot(any) != 0ull;
        };
        int o = next_present(-1);
        int scur[CB], snext[CB];
        double bcur[CB][KC], bnext[CB][KC];
        bool has_cur = false, has_next = false;
        if (o < kNumOffsets) {
            const AT* src = Asw + (size_t)o * NKC * CH;
            for (int e = tid; e < CH; e += NT) abuf[0][e] = src[e];
            has_cur = load_src(o, scur);
            if (has_cur) load_b(scur, 0, bcur);
            has_next = load_src(next_present(o), snext);
        }
        __syncthreads();
        int buf = 0, c = 0;
        while (o < kNumOffsets) {
            int on = o, cn = c + 1;
            if (cn == NKC) {
                cn = 0;
                on = next_present(o);
            }
            AT pre[PER];
            if (on < kNumOffsets) {
                const AT* src = Asw + ((size_t)on * NKC + cn) * CH;
#pragma unroll
                for (int i = 0; i < PER; ++i) {
                    const int e = tid + NT * i;
                    pre[i] = (e < CH) ? src[e] : AT{};
                }
            }
            // sources and B of the next item
            const bool nxt_has = (cn == 0) ? has_next : has_cur;
            if (on < kNumOffsets && nxt_has) {
                if (cn == 0) load_b(snext, 0, bnext);
                else load_b(scur, cn, bnext);
            }
            if (has_cur) {
                const AT* ab = abuf[buf] + lane;
                if constexpr (MIX) {
                    float bh[CB][KC], bl[CB][KC];
#pragma unroll
                    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
                        for (int kk = 0; kk < KC; ++kk) {
                            bh[cb][kk] = (float)bcur[cb][kk];
                            bl[cb][kk] = (float)(bcur[cb][kk] - (double)bh[cb][kk]);
                        }
#pragma unroll
                    for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                        for (int rb = 0; rb < RB; ++rb) {
                            const float2 a = ab[(rb * KC + kk) * 64];
#pragma unroll
                            for (int cb = 0; cb < CB; ++cb) {
                                acc[cb][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, bh[cb][kk], acc[cb][rb], 0, 0, 0);
                                acc[cb][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, bl[cb][kk], acc[cb][rb], 0, 0, 0);
                                acc[cb][rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, bh[cb][kk], acc[cb][rb], 0, 0, 0);
                            }
                        }
                } else {
#pragma unroll
                    for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                        for (int rb = 0; rb < RB; ++rb) {
                            const double a = ab[(rb * KC + kk) * 64];
#pragma unroll
                            for (int cb = 0; cb < CB; ++cb)
                                acc[cb][rb] =
                                    __builtin_amdgcn_mfma_f64_16x16x4f64(a, bcur[cb][kk], acc[cb][rb], 0, 0, 0);
                        }
                }
            }
            if (on < kNumOffsets) {
#pragma unroll
                for (int i = 0; i < PER; ++i) {
                    const int e = tid + NT * i;
                    if (e < CH) abuf[buf ^ 1][e] = pre[i];
                }
            }
            __syncthreads();
            buf ^= 1;
            if (cn == 0 && on < kNumOffsets) {
#pragma unroll
                for (int cb = 0; cb < CB; ++cb) scur[cb] = snext[cb];
                has_cur = has_next;
                has_next = load_src(next_present(on), snext);
            }
#pragma unroll
            for (int cb = 0; cb < CB; ++cb)
#pragma unroll
                for (int kk = 0; kk < KC; ++kk) bcur[cb][kk] = bnext[cb][kk];
            o = on;
            c = cn;
        }
    }
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
        if (rank[cb] >= nb) continue;
        const int t = order[rank[cb]];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // D(i, j) of lane j + 16 kq, register r: i = kq + 4r (f64), 4kq + r (f32)
                const int row = MIX ? 16 * rb + 4 * kq + r : 16 * rb + kq + 4 * r;
                if (row < S) L[(size_t)t * S + row] = (double)acc[cb][rb][r];
            }
    }
}

// ---- M2M / L2L on FP64 matrix cores ----------------------------------------------
// out[t] (+)= sum_o T[o] in[src(o, t)] over the 8 octant operators for 16
// targets of one level per workgroup of 8 waves: wave o applies octant o's
// operator (v_mfma_f64_16x16x4_f64, all A fragments and B values of a K-chunk
// loaded before its MFMAs), and the 8 partial 16 x S tiles are summed in LDS
// in octant order (deterministic).
//   MODE 0 (M2M, FmmExpansions3D.hpp:242-265): targets are the level's
//     non-leaf boxes, src(o, t) = child of t in octant o; overwrites Mhat[t].
//   MODE 1 (L2L, :307-330): targets are the level's boxes in octant order,
//     src(o, t) = parent if octant(t) == o (one operator per target); adds
//     into Lhat[t].
constexpr int kLevelWaves = 8;
template <int P, int MODE>
__global__ void __launch_bounds__(kLevelWaves * 64) k_level_mfma(int n, const int* __restrict__ tgt,
                                                                const int* __restrict__ child_first,
                                                                const unsigned char* __restrict__ mask,
                                                                const int* __restrict__ parent,
                                                                const int* __restrict__ octant,
                                                                const double* __restrict__ Tsw,
                                                                const double* __restrict__ in,
                                                                double* __restrict__ out) {
    constexpr int S = nterms(P);
    constexpr int RB = m2l_rb(S), NKC = m2l_nkc(S), KC = m2l_kc(S), CH = m2l_chunk(S);
    __shared__ double red[16 * RB * 16];  // [row][target], rows padded to RB*16
    const int tid = threadIdx.x, o = tid >> 6, lane = tid & 63, j = lane & 15, kq = lane >> 4;
    const int ti = blockIdx.x * 16 + j;
    const bool valid = ti < n;
    const int t = valid ? tgt[ti] : -1;
    int src = -1;
    if (valid) {
        if (MODE == 0) {
            const unsigned mk = mask[t];
            if ((mk >> o) & 1u) src = child_first[t] + __popc(mk & ((1u << o) - 1u));
        } else if (octant[t] == o) {
            src = parent[t];
        }
    }
    for (int e = tid; e < 16 * RB * 16; e += kLevelWaves * 64) red[e] = 0.0;
    dbl4 acc[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) acc[rb] = dbl4{0.0, 0.0, 0.0, 0.0};
    const bool active = __ballot(src >= 0) != 0ull;
    if (active) {
        const double* __restrict__ ms = in + (size_t)(src >= 0 ? src : 0) * S;
        for (int c = 0; c < NKC; ++c) {
            double bv[KC], av[RB * KC];
            const double* __restrict__ A = Tsw + ((size_t)o * NKC + c) * CH + lane;
#pragma unroll
            for (int e = 0; e < RB * KC; ++e) av[e] = A[e * 64];
#pragma unroll
            for (int kk = 0; kk < KC; ++kk) {
                const int k = 4 * (c * KC + kk) + kq;
                bv[kk] = (src >= 0 && k < S) ? ms[k] : 0.0;
            }
#pragma unroll
            for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
                    acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[rb * KC + kk], bv[kk], acc[rb], 0, 0, 0);
        }
    }
    // octant-ordered sum of the partial tiles
    for (int w = 0; w < kLevelWaves; ++w) {
        __syncthreads();
        if (w == o && active) {
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                for (int r = 0; r < 4; ++r) red[(16 * rb + kq + 4 * r) * 16 + j] += acc[rb][r];
        }
    }
    __syncthreads();
    // write: thread e covers (row, target) pairs
    for (int e = tid; e < S * 16; e += kLevelWaves * 64) {
        const int row = e >> 4, jj = e & 15;
        const int tj = blockIdx.x * 16 + jj;
        if (tj >= n) continue;
        const int tt = tgt[tj];
        if (MODE == 0) out[(size_t)tt * S + row] = red[row * 16 + jj];
        else out[(size_t)tt * S + row] += red[row * 16 + jj];
    }
}

// ---- L2P: one thread per sorted body (far field from its leaf's local).
template <int P>
__global__ void __launch_bounds__(kT) k_far(TreeView t, const double* __restrict__ L, double* __restrict__ far,
                                            double* __restrict__ fpot) {
    constexpr int S = nterms(P);
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= t.n) return;
    const int b = t.leaf_of[i];
    const double xi = t.x[i], yi = t.y[i], zi = t.z[i];
    auto dhat = [&](int box, double& dx, double& dy, double& dz) {
        const double* c = t.cdd + 6 * (size_t)box;
        const double ih = t.invh[box];
        dx = offset_from(xi, Centre{c[0], c[1]}) * ih;
        dy = offset_from(yi, Centre{c[2], c[3]}) * ih;
        dz = offset_from(zi, Centre{c[4], c[5]}) * ih;
    };
    double dx, dy, dz, gx, gy, gz;
    dhat(b, dx, dy, dz);
    const double* lb = L + (size_t)b * S;
    (void)S;
    local_grad<P>([&](int k) { return lb[k]; }, dx, dy, dz, gx, gy, gz);
    const double s2 = t.invh[b] * t.invh[b];
    far[3 * (size_t)i] = gx * s2;
    far[3 * (size_t)i + 1] = gy * s2;
    far[3 * (size_t)i + 2] = gz * s2;
    if (fpot) fpot[i] = local_pot<P>([&](int k) { return lb[k]; }, dx, dy, dz) * t.invh[b];
}

// ---- P2P over U + final combine (AdaptiveFMM3D.hpp:96-106) ------------------
// One wave per target leaf. The source bodies of U(b) (contiguous ranges in
// sorted order) are streamed as one concatenated sequence in full 64-body LDS
// tiles. The wave is split into G = floor(64/T) lane groups of T = the leaf's
// body count (at least 8, at most 64); group g takes the tile's sources
// g, g+G, ... so short leaves still keep all lanes busy. Partial sums are
// reduced across groups in a fixed order (deterministic). The kernel is
// (|d|^2 + eps^2)^{-3/2} from v_rsq_f64 refined by two Newton steps; the self
// pair is skipped by index, as in the reference.
constexpr int kP2PMaxU = 256;

__device__ __forceinline__ double rsqrt_f64(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = y * (1.5 - hx * y * y);
    y = y * (1.5 - hx * y * y);
    return y;
}

// MIX: the pair difference stays FP64 (it carries the cancellation), then
// r^2, v_rsq_f32 and the pair term run in FP32; each lane sums its terms of
// one 64-body source tile in FP32 and adds that partial to its FP64 accumulator.
template <bool MIX, bool POT>
__global__ void __launch_bounds__(64) k_p2p(TreeView t, const double* __restrict__ far, double* __restrict__ acc,
                                           const double* __restrict__ fpot, double* __restrict__ phi) {
    __shared__ double sx[64], sy[64], sz[64], sm[64];
    __shared__ int sj[64];
    __shared__ int ufirst[kP2PMaxU], ucount[kP2PMaxU];
    __shared__ double red[4][64];
    const int li = blockIdx.x;
    const int b = t.leaf_ids[li];
    const int f = t.seg_first[b], nt = t.seg_count[b];
    const int lane = threadIdx.x;
    const int q = t.bsys[b];
    const double G = t.sys_G[q], eps2 = t.sys_eps2[q];
    const int ua = t.u_off[li], nu = t.u_off[li + 1] - ua;
    const bool cached = nu <= kP2PMaxU;
    if (cached)
        for (int e = lane; e < nu; e += 64) {
            const int c = t.u_idx[ua + e];
            ufirst[e] = t.seg_first[c];
            ucount[e] = t.seg_count[c];
        }
    __syncthreads();
    auto range = [&](int e, int& first, int& cnt) {
        if (cached) {
            first = ufirst[e];
            cnt = ucount[e];
        } else {
            const int c = t.u_idx[ua + e];
            first = t.seg_first[c];
            cnt = t.seg_count[c];
        }
    };
    const int T = nt > 64 ? 64 : (nt < 8 ? 8 : nt);
    const int Gs = 64 / T;
    const int g = lane / T, tl = lane % T;
    const int kfirst = g < Gs ? g : 64;  // lanes past the last whole group idle
    for (int tb = 0; tb < nt; tb += T) {
        const int i = f + tb + tl;
        const bool act = g < Gs && tb + tl < nt;
        const double xi = act ? t.x[i] : 0.0, yi = act ? t.y[i] : 0.0, zi = act ? t.z[i] : 0.0;
        double ax = 0.0, ay = 0.0, az = 0.0, ap = 0.0;
        int e = 0, off = 0;  // stream position: range e, offset off
        while (e < nu) {
            // lane takes stream position (e, off) + lane
            int ee = e, oo = off + lane, fst = 0, cnt = 0;
            bool ok = false;
            while (ee < nu) {
                range(ee, fst, cnt);
                if (oo < cnt) {
                    ok = true;
                    break;
                }
                oo -= cnt;
                ++ee;
            }
            if (ok) {
                const int j = fst + oo;
                sx[lane] = t.x[j];
                sy[lane] = t.y[j];
                sz[lane] = t.z[j];
                sm[lane] = t.m[j];
                sj[lane] = j;
            } else {
                sj[lane] = -1;
            }
            // advance the stream by 64 bodies (wave-uniform)
            {
                int rem = 64;
                while (e < nu) {
                    int f2, c2;
                    range(e, f2, c2);
                    if (off + rem < c2) {
                        off += rem;
                        rem = 0;
                        break;
                    }
                    rem -= c2 - off;
                    off = 0;
                    ++e;
                }
            }
            __syncthreads();
            if constexpr (MIX) {
                const float e2 = (float)eps2;
                float fx = 0.f, fy = 0.f, fz = 0.f, fp = 0.f;
                for (int k = kfirst; k < 64; k += Gs) {
                    const int j = sj[k];
                    if (j < 0) break;
                    const float dx = (float)(sx[k] - xi), dy = (float)(sy[k] - yi), dz = (float)(sz[k] - zi);
                    const float r2 = dx * dx + dy * dy + dz * dz + e2;
                    const float rin = __builtin_amdgcn_rsqf(r2);
                    // the self pair is selected away, not multiplied by zero
                    // (with eps = 0 its 1/r is infinite)
                    const float mj = (float)sm[k];
                    const float s = (j != i) ? mj * (rin * rin * rin) : 0.f;
                    fx += s * dx;
                    fy += s * dy;
                    fz += s * dz;
                    if constexpr (POT) fp += (j != i) ? mj * rin : 0.f;
                }
                ax += (double)fx;
                ay += (double)fy;
                az += (double)fz;
                ap += (double)fp;
            } else {
                for (int k = kfirst; k < 64; k += Gs) {
                    const int j = sj[k];
                    if (j < 0) break;
                    const double dx = sx[k] - xi, dy = sy[k] - yi, dz = sz[k] - zi;
                    const double r2 = dx * dx + dy * dy + dz * dz + eps2;
                    const double rin = rsqrt_f64(r2);
                    const double s = (j != i) ? sm[k] * (rin * rin * rin) : 0.0;
                    ax += s * dx;
                    ay += s * dy;
                    az += s * dz;
                    if constexpr (POT) ap += (j != i) ? sm[k] * rin : 0.0;
                }
            }
            __syncthreads();
        }
        if (Gs > 1) {
            red[0][lane] = ax;
            red[1][lane] = ay;
            red[2][lane] = az;
            if constexpr (POT) red[3][lane] = ap;
            __syncthreads();
            if (g == 0) {
                for (int q = 1; q < Gs; ++q) {
                    ax += red[0][q * T + tl];
                    ay += red[1][q * T + tl];
                    az += red[2][q * T + tl];
                    if constexpr (POT) ap += red[3][q * T + tl];
                }
            }
            __syncthreads();
        }
        if (act && g == 0) {
            const int o = t.orig[i];
            acc[3 * (size_t)o] = G * (far[3 * (size_t)i] + ax);
            acc[3 * (size_t)o + 1] = G * (far[3 * (size_t)i + 1] + ay);
            acc[3 * (size_t)o + 2] = G * (far[3 * (size_t)i + 2] + az);
            if constexpr (POT) phi[o] = -G * (fpot[i] + ap);
        }
    }
}

template <int P>
void run(Pool& pool, hipStream_t s, const TreeView& t, const TreeMeta& meta, const FmmTables& tab, bool mixed,
         double* acc, double* phi, hipEvent_t* ev) {
    constexpr int S = nterms(P);
    const int nb = t.nb;
    double* M = pool.get<double>("Mexp", (size_t)nb * S);
    double* L = pool.get<double>("Lexp", (size_t)nb * S);
    double* far = pool.get<double>("far", 3 * (size_t)t.n);
    double* fpot = phi ? pool.get<double>("farpot", t.n) : nullptr;
    if (ev) AFMM_CHECK(hipEventRecord(ev[0], s));
    hipLaunchKernelGGL((k_accumulate<P, 0>), dim3(t.nleaves), dim3(64), 0, s, t, M);
    for (int l = t.max_level - 1; l >= 0; --l) {
        const int off = meta.nonleaf_off[l], cnt = meta.nonleaf_off[l + 1] - off;
        if (cnt > 0)
            hipLaunchKernelGGL((k_level_mfma<P, 0>), dim3((cnt + 15) / 16), dim3(kLevelWaves * 64), 0, s, cnt,
                               t.nonleaf_ids + off,
                               t.child_first, t.child_mask, t.parent, t.octant, tab.m2m_sw, M, M);
    }
    if (ev) AFMM_CHECK(hipEventRecord(ev[1], s));
    {
        // M2L targets in (level, octant) order; transposed source table; tile plan
        constexpr int TILE = m2l_tile<P>();
        const int np = (nb + TILE - 1) / TILE * TILE;
        const int nwg = np / TILE;
        int* rank = pool.get<int>("m2l_rank", nb);
        int* srcT = pool.get<int>("m2l_srcT", (size_t)kNumOffsets * np);
        unsigned char* present = pool.get<unsigned char>("m2l_present", (size_t)nwg * kNumOffsets);
        hipLaunchKernelGGL(k_rank_of, dim3(blocks(nb)), dim3(kT), 0, s, nb, t.lo_order, rank);
        hipLaunchKernelGGL(k_fill_int, dim3(blocks((long long)kNumOffsets * np)), dim3(kT), 0, s, srcT,
                           (size_t)kNumOffsets * np, -1);
        hipLaunchKernelGGL(k_m2l_src, dim3(blocks(nb)), dim3(kT), 0, s, nb, np, rank, t.vl, t.nv, t.ic, srcT);
        hipLaunchKernelGGL(k_m2l_plan, dim3(nwg * kNumOffsets), dim3(64), 0, s, TILE, np, nb, srcT, present);
        if (ev) AFMM_CHECK(hipEventRecord(ev[5], s));
        if (mixed)
            hipLaunchKernelGGL((k_m2l_mfma<P, true>), dim3(nwg), dim3(kM2LWaves * 64), 0, s, nb, np, nb, t.lo_order,
                               srcT, present, tab.m2l_hl, M, L);
        else
            hipLaunchKernelGGL((k_m2l_mfma<P, false>), dim3(nwg), dim3(kM2LWaves * 64), 0, s, nb, np, nb, t.lo_order,
                               srcT, present, tab.m2l_sw, M, L);
    }
    if (ev) AFMM_CHECK(hipEventRecord(ev[2], s));
    if (t.n_x > 0) hipLaunchKernelGGL((k_accumulate<P, 1>), dim3(nb), dim3(64), 0, s, t, L);
    for (int l = 1; l <= t.max_level; ++l) {
        const int off = meta.level_off[l], cnt = meta.level_off[l + 1] - off;
        hipLaunchKernelGGL((k_level_mfma<P, 1>), dim3((cnt + 15) / 16), dim3(kLevelWaves * 64), 0, s, cnt,
                           t.lo_order + off,
                           t.child_first, t.child_mask, t.parent, t.octant, tab.l2l_sw, L, L);
    }
    if (ev) AFMM_CHECK(hipEventRecord(ev[3], s));
    hipLaunchKernelGGL(k_far<P>, dim3(blocks(t.n)), dim3(kT), 0, s, t, L, far, fpot);
    if (t.n_w > 0) {
        if (fpot) hipLaunchKernelGGL((k_m2p<P, true>), dim3(t.nleaves), dim3(64), 0, s, t, M, far, fpot);
        else hipLaunchKernelGGL((k_m2p<P, false>), dim3(t.nleaves), dim3(64), 0, s, t, M, far, fpot);
    }
    const dim3 pg(t.nleaves), pb(64);
    if (mixed && phi)
        hipLaunchKernelGGL((k_p2p<true, true>), pg, pb, 0, s, t, far, acc, fpot, phi);
    else if (mixed)
        hipLaunchKernelGGL((k_p2p<true, false>), pg, pb, 0, s, t, far, acc, fpot, phi);
    else if (phi)
        hipLaunchKernelGGL((k_p2p<false, true>), pg, pb, 0, s, t, far, acc, fpot, phi);
    else
        hipLaunchKernelGGL((k_p2p<false, false>), pg, pb, 0, s, t, far, acc, fpot, phi);
    if (ev) AFMM_CHECK(hipEventRecord(ev[4], s));
    AFMM_CHECK(hipGetLastError());
}

}  // namespace

void run_fmm(Pool& pool, hipStream_t s, const TreeView& tv, const TreeMeta& meta, int p, const FmmTables& tab,
             bool mixed, double* d_acc, double* d_phi, hipEvent_t* ev) {
    if (mixed && !tab.m2l_hl) throw std::runtime_error("split FP32 M2L operators not loaded");
    switch (p) {
#define AFMM_CASE(PP) \
    case PP: run<PP>(pool, s, tv, meta, tab, mixed, d_acc, d_phi, ev); return;
        AFMM_CASE(1) AFMM_CASE(2) AFMM_CASE(3) AFMM_CASE(4) AFMM_CASE(5) AFMM_CASE(6)
        AFMM_CASE(7) AFMM_CASE(8) AFMM_CASE(9) AFMM_CASE(10) AFMM_CASE(11) AFMM_CASE(12)
#undef AFMM_CASE
    }
    throw std::runtime_error("unsupported expansion order");
}

}  // namespace afmm
