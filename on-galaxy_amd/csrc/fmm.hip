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
#include <type_traits>

#include "afmm_internal.h"
#include "afmm_math.h"

namespace afmm {
namespace {

constexpr int kT = 256;
inline int blocks(long long n, int t = kT) { return (int)std::max<long long>(1, (n + t - 1) / t); }

// ---- XCD-aware workgroup order ------------------------------------------------
// MI355X deals workgroups round-robin to its 8 XCDs, each with its own L2.
// Kernels whose neighbouring work items (tiles or leaves in tree order) read
// the same neighbour data launch xcd_grid(n) workgroups and take item
// xcd_item(): each XCD then runs one contiguous range of items, in order, and
// the shared reads hit its L2. Items >= n return at once.
// Per-leaf kernels, whose cost varies from leaf to leaf in spatial clusters,
// use chunks of K items instead (xcd_chunk_*): chunks go round-robin to the
// XCDs, which keeps each XCD's share balanced and its reads local within a chunk.
constexpr int kXcds = 8;
inline int xcd_grid(int n) { return (n + kXcds - 1) / kXcds * kXcds; }
__device__ __forceinline__ int xcd_item() {
    return (int)(blockIdx.x % kXcds) * (int)(gridDim.x / kXcds) + (int)(blockIdx.x / kXcds);
}
constexpr int kLeafChunk = 32;
inline int xcd_chunk_grid(int n) { return (n + kXcds * kLeafChunk - 1) / (kXcds * kLeafChunk) * (kXcds * kLeafChunk); }
__device__ __forceinline__ int xcd_chunk_item() {
    const int x = (int)(blockIdx.x % kXcds), slot = (int)(blockIdx.x / kXcds);
    return ((slot / kLeafChunk) * kXcds + x) * kLeafChunk + slot % kLeafChunk;
}

// ---- wave helpers -------------------------------------------------------------
// Sums each of a lane's H register values over the wave by recursive halving:
// at mask M = 32, 16, ..., 1 a lane keeps the lower ceil(H/2) of its values
// when (lane & M) == 0, else the upper ones, and adds its partner's copy of
// them. Every coefficient's total is one fixed sum tree (deterministic), and a
// wave exchanges about H values instead of 6 H for H separate butterflies. On
// return lane l holds the totals of coefficients first .. first + cnt - 1 in
// v[0 .. cnt) (cnt <= 2 for H <= 128).
template <int H, int M>
__device__ __forceinline__ void wave_halving_sum(double* v, int lane, int& first, int& cnt) {
    if constexpr (M >= 1) {
        constexpr int H2 = (H + 1) / 2;
        const bool up = (lane & M) != 0;
#pragma unroll
        for (int q = 0; q < H2; ++q) {
            const double lo = v[q];
            const double hi = (H2 + q < H) ? v[H2 + q] : 0.0;
            const double keep = up ? hi : lo;
            v[q] = keep + __shfl_xor(up ? lo : hi, M, 64);
        }
        if (up) {
            first += H2;
            cnt = cnt > H2 ? cnt - H2 : 0;
        } else {
            cnt = cnt < H2 ? cnt : H2;
        }
        wave_halving_sum<H2, M / 2>(v, lane, first, cnt);
    }
}

// ---- P2M and P2L: per-body coefficient accumulation -------------------------------
// One wave per target box. Lanes take the source bodies 64 at a time from the
// packed stream; each lane accumulates q conj(R(d)) (P2M, FmmExpansions3D.hpp:
// 148-163) or q conj(S(d)) (P2L, AdaptiveFMM3D.hpp:117-128) of its bodies into
// a register window of the S coefficients; the windows are then summed across
// the wave by recursive halving (fixed order). Orders above 8 take two windows
// (two passes over the stream) to bound registers.
//   KIND 0: targets = leaves, sources = the leaf's own bodies, Mhat = result.
//   KIND 1: targets = boxes, sources = bodies of X(box) leaves, Lhat += result.
//   KIND 2: as 1 into a separate array Lx (k_add_x then adds it to Lhat after
//   M2L: the same sums in the same order, and free to run beside M2L).
template <int S>
constexpr int coef_windows() { return S <= 96 ? 1 : ((S + 127) / 128 > 2 ? (S + 127) / 128 : 2); }
template <int S>
constexpr int coef_window() { return (S + coef_windows<S>() - 1) / coef_windows<S>(); }

// Bodies of a list of boxes staged through LDS in chunks of kStreamChunk list
// entries: their (first body, count) and the running body offset, so that
// stream position q maps to its body by a binary search in LDS rather than by
// walking the list in global memory.
constexpr int kStreamChunk = 256;
struct StreamLds {
    int first[kStreamChunk];
    int pre[kStreamChunk + 1];
};

// Loads entries [e0, e0 + m) of `list` (m <= kStreamChunk) into LDS and
// returns the number of bodies they hold (one-wave workgroup).
__device__ __forceinline__ int stage_chunk(StreamLds& sl, const int* list, int e0, int m, const int* seg_first,
                                           const int* seg_count, int lane) {
    {
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
        int incl = run;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(incl, d, 64);
            if (lane >= d) incl += t;
        }
        const int base = incl - run;
#pragma unroll
        for (int r = 0; r < kStreamChunk / 64; ++r) sl.pre[lane * (kStreamChunk / 64) + r] = base + cnt[r];
        if (lane == 63) sl.pre[kStreamChunk] = incl;
    }
    __syncthreads();
    return sl.pre[kStreamChunk];
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

// Coefficients C0 .. C0 + CW - 1 of one target box: lanes take the staged
// bodies 64 at a time, then the wave sums its windows by recursive halving.
// P2M (KIND 0) reads the leaf's own contiguous bodies f0 .. f0 + n0 - 1
// directly; P2L streams the bodies of the X list through LDS.
template <int P, int KIND, int C0, int CW>
__device__ __forceinline__ void accumulate_window(const TreeView& t, const int* list, int nlist, int f0, int n0,
                                                  const double* ch, double ih, int lane, StreamLds& sl,
                                                  double* __restrict__ out, size_t out_base) {
    constexpr int S = nterms(P);
    double acc[CW];
#pragma unroll
    for (int k = 0; k < CW; ++k) acc[k] = 0.0;
    for (int e0 = 0; e0 < nlist; e0 += kStreamChunk) {
        const int m = min(kStreamChunk, nlist - e0);
        int total = n0;
        if constexpr (KIND != 0) {
            __syncthreads();  // previous chunk's readers are done
            total = stage_chunk(sl, list, e0, m, t.seg_first, t.seg_count, lane);
        }
        // The next slot's body is located and loaded before the current slot's
        // harmonics, so its search and load latency overlap the arithmetic.
        double nx = 0.0, ny = 0.0, nz = 0.0, nm = 0.0;
        auto fetch = [&](int q) {
            if (q < total) {
                const int i = KIND == 0 ? f0 + q : stream_body(sl, m, q);
                nx = t.x[i], ny = t.y[i], nz = t.z[i], nm = t.m[i];
            }
        };
        fetch(lane);
        for (int q0 = 0; q0 < total; q0 += 64) {
            const bool valid = q0 + lane < total;
            const double bx = nx, by = ny, bz = nz, bm = nm;
            fetch(q0 + 64 + lane);
            if (valid) {
                const double dx = offset_from(bx, Centre{ch[0], ch[1]}) * ih;
                const double dy = offset_from(by, Centre{ch[2], ch[3]}) * ih;
                const double dz = offset_from(bz, Centre{ch[4], ch[5]}) * ih;
                auto add = [&](int k, double v) {
                    if (k >= C0 && k < C0 + CW) acc[k - C0] += v;
                };
                if (KIND == 0) p2m_point<P>(bm, dx, dy, dz, add);
                else p2l_point<P>(bm, dx, dy, dz, add);
            }
        }
    }
    static_assert(CW >= 2 && CW <= 128, "halving leaves at most two totals per lane");
    int first = 0, cnt = CW;
    wave_halving_sum<CW, 32>(acc, lane, first, cnt);
#pragma unroll
    for (int k = 0; k < 2; ++k)
        if (k < cnt && C0 + first + k < S) {
            if constexpr (KIND == 1) out[out_base + C0 + first + k] += acc[k];
            else out[out_base + C0 + first + k] = acc[k];
        }
}

// One wave per box (four waves per box with X lists measured 2.7x slower:
// most X ranges hold too few bodies to occupy them).
template <int P, int KIND>
__global__ void __launch_bounds__(64) k_accumulate(TreeView t, double* __restrict__ out) {
    constexpr int S = nterms(P);
    constexpr int CW = coef_window<S>();
    __shared__ StreamLds sl;
    const int lane = threadIdx.x;
    int b, nlist;
    const int* list;
    const int item = xcd_chunk_item();
    if (KIND == 0) {
        if (item >= t.nleaves) return;
        b = t.leaf_ids[item];
        list = t.leaf_ids + item;
        nlist = 1;
    } else {
        if (item >= t.nb) return;
        b = item;
        const int xa = t.x_off[b];
        list = t.x_idx + xa;
        nlist = t.x_off[b + 1] - xa;
        if (nlist == 0) return;
    }
    const double ch[6] = {t.cdd[6 * (size_t)b], t.cdd[6 * (size_t)b + 1], t.cdd[6 * (size_t)b + 2],
                          t.cdd[6 * (size_t)b + 3], t.cdd[6 * (size_t)b + 4], t.cdd[6 * (size_t)b + 5]};
    const double ih = t.invh[b];
    const int f0 = KIND == 0 ? t.seg_first[b] : 0, n0 = KIND == 0 ? t.seg_count[b] : 0;
    accumulate_window<P, KIND, 0, CW>(t, list, nlist, f0, n0, ch, ih, lane, sl, out, (size_t)b * S);
    if constexpr (CW < S)
        accumulate_window<P, KIND, CW, (2 * CW < S ? CW : S - CW)>(t, list, nlist, f0, n0, ch, ih, lane, sl, out,
                                                                   (size_t)b * S);
    if constexpr (2 * CW < S)
        accumulate_window<P, KIND, 2 * CW, S - 2 * CW>(t, list, nlist, f0, n0, ch, ih, lane, sl, out, (size_t)b * S);
    static_assert(3 * CW >= S, "three coefficient windows cover every order built");
}

// ---- M2P over W (AdaptiveFMM3D.hpp:95, FmmExpansions3D.hpp:180-200) --------------
// One workgroup of kM2PWaves waves per target leaf. Its L = 64 * kM2PWaves lanes
// form G = L / T lane groups of T = max(n_leaf, L / m2p_groups) lanes (T = L
// above L bodies); group g handles W entries g, g + G, ... with their
// multipoles staged in LDS. Partial fields are reduced across groups in a
// fixed order and added to far[] (and, with POT, the potential to fpot[]).
// Four waves per leaf rather than one: the W lists are heavy-tailed (a few
// clustered leaves hold 200-300 entries), and one wave per leaf left those
// waves running long after the rest of the grid had drained.
// At p <= 8 the kernel is held to 128 VGPRs (four waves per SIMD; a few bytes
// spill), measured 5% faster than 143 VGPRs at three waves; above p = 8 that
// limit spills hundreds of bytes and is not applied.
constexpr int kM2PWaves = 4;
template <int P>
constexpr int m2p_groups() { return P <= 8 ? 32 : 16; }
template <int P>
constexpr int m2p_min_waves() { return P <= 8 ? 4 : 1; }

// ADD: far += the field (after L2P wrote far); otherwise far = the field, zero
// for leaves without W (the separate array L2P adds later; the same sums).
// order: the leaves by descending work (m2p_order); workgroup k takes leaf
// order[k], which changes when a leaf runs, not what it sums.
template <int P, bool POT, bool ADD>
__global__ void __launch_bounds__(kM2PWaves * 64, m2p_min_waves<P>()) k_m2p(TreeView t, const double* __restrict__ M,
                                                       double* __restrict__ far, double* __restrict__ fpot,
                                                       const int* __restrict__ order) {
    constexpr int S = nterms(P), L = kM2PWaves * 64, GM = m2p_groups<P>();
    __shared__ double mw[GM][S];
    __shared__ double red[4][L];
    const int item = xcd_chunk_item();
    if (item >= t.nleaves) return;
    const int li = order[item];
    const int wa = t.w_off[li], nw = t.w_off[li + 1] - wa;
    const int b = t.leaf_ids[li];
    const int f = t.seg_first[b], nt = t.seg_count[b];
    const int tid = threadIdx.x;
    if (nw == 0) {  // no W: a zero field, so that L2P can add farw everywhere
        if constexpr (ADD) return;
        for (int k = tid; k < nt; k += kM2PWaves * 64) {
            const size_t i = (size_t)(f + k);
            far[3 * i] = 0.0;
            far[3 * i + 1] = 0.0;
            far[3 * i + 2] = 0.0;
            if constexpr (POT) fpot[i] = 0.0;
        }
        return;
    }
    const int T = nt > L ? L : max(nt, L / GM);
    const int Gs = L / T;  // <= GM
    const int g = tid / T, tl = tid % T;
    for (int tb = 0; tb < nt; tb += T) {
        const int i = f + tb + tl;
        const bool act = g < Gs && tb + tl < nt;
        const double xi = act ? t.x[i] : 0.0, yi = act ? t.y[i] : 0.0, zi = act ? t.z[i] : 0.0;
        double fx = 0.0, fy = 0.0, fz = 0.0, fp = 0.0;
        for (int r = 0; r < nw; r += Gs) {
            __syncthreads();
            const int nq = min(Gs, nw - r);
            for (int e = tid; e < nq * S; e += L) {
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
                double gx, gy, gz, pt = 0.0;
                const double* mg = mw[g];
                multipole_grad<P, POT>([&](int k) { return mg[k]; }, dx, dy, dz, gx, gy, gz, &pt);
                const double s2 = ih * ih;
                fx += gx * s2;
                fy += gy * s2;
                fz += gz * s2;
                if constexpr (POT) fp += pt * ih;
            }
        }
        if (Gs > 1) {
            __syncthreads();
            red[0][tid] = fx;
            red[1][tid] = fy;
            red[2][tid] = fz;
            red[3][tid] = fp;
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
            if constexpr (ADD) {
                far[3 * (size_t)i] += fx;
                far[3 * (size_t)i + 1] += fy;
                far[3 * (size_t)i + 2] += fz;
                if constexpr (POT) fpot[i] += fp;
            } else {
                far[3 * (size_t)i] = fx;
                far[3 * (size_t)i + 1] = fy;
                far[3 * (size_t)i + 2] = fz;
                if constexpr (POT) fpot[i] = fp;
            }
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

__global__ void __launch_bounds__(kT) k_zero_int(int* a, size_t n) {
    const size_t i = blockIdx.x * (size_t)kT + threadIdx.x;
    if (i < n) a[i] = 0;
}

// The offset slots of afmm_math.h for device code (srcT rows, presence and
// operators are all kept in slot order).
__constant__ M2LSlots kM2LSlots = make_m2l_slots();

// Source-table rows: a target at integer coordinate 2 P + c (c = 0 or 1 per
// axis, the parities par = 4 cx + 2 cy + cz) meets its V-list sources, the
// children of its parent's colleagues P - 1 .. P + 1 that are not adjacent,
// only at offsets d = ic(target) - ic(source) in [c - 3, c + 2]: 6 of the 7
// values per axis, 216 of the 343 offsets. Row of offset index oi (of
// offset_index) for parities par, or -1 where the target cannot meet it.
constexpr int kSrcRows = 216;
__device__ __forceinline__ int src_row(int oi, int par) {
    const int ux = oi / 49 - (par >> 2), uy = (oi / 7) % 7 - ((par >> 1) & 1), uz = oi % 7 - (par & 1);
    return (unsigned)ux < 6u && (unsigned)uy < 6u && (unsigned)uz < 6u ? (ux * 6 + uy) * 6 + uz : -1;
}
__device__ __forceinline__ int ic_parity(const int* __restrict__ ic, int b) {
    return ((ic[3 * (size_t)b] & 1) << 2) | ((ic[3 * (size_t)b + 1] & 1) << 1) | (ic[3 * (size_t)b + 2] & 1);
}

// srcT[row * np + r] = the source box of target order[r] at the offset of that
// row (src_row for the target's parities), -1 where it has none, and
// present[tile][slot] = 1 when some target of the tile has a source at that
// slot (present zeroed before). One block per 16 consecutive targets of the
// M2L order: their V lists fill a rows x 16 table in LDS (initially -1), which
// is then written out row by row, 64 contiguous bytes per row, so the whole
// table is written once and coalesced (no separate fill of -1). Presence: one
// integer atomic per (slot, run of the block's targets in one tile) with a
// source there (order-free; a block's targets span one tile or a few small
// ones).
constexpr int kSrcTargets = 16;
__global__ void __launch_bounds__(kT) k_m2l_src(int nb, int np, const int* __restrict__ order,
                                               const int* __restrict__ vl, const int* __restrict__ nv,
                                               const int* __restrict__ ic, int* __restrict__ srcT,
                                               const int* __restrict__ tstart, int ntiles,
                                               int* __restrict__ present) {
    __shared__ int tab[kSrcRows][kSrcTargets];
    __shared__ int box[kSrcTargets], cnt[kSrcTargets], tile_of[kSrcTargets], par[kSrcTargets];
    // consecutive blocks on one XCD, so the two 64-byte halves of a 128-byte
    // row segment are written through the same L2
    const int r0 = xcd_item() * kSrcTargets, tid = threadIdx.x;
    if (r0 >= nb) return;
    const int nt = min(kSrcTargets, nb - r0);
    for (int e = tid; e < kSrcRows * kSrcTargets; e += kT) (&tab[0][0])[e] = -1;
    if (tid < nt) {
        box[tid] = order[r0 + tid];
        cnt[tid] = nv[box[tid]];
        par[tid] = ic_parity(ic, box[tid]);
        // the tile of the target: the last k with tstart[k] <= rank (one
        // search per target, side by side)
        int lo = 0, hi = ntiles - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tstart[mid] <= r0 + tid) lo = mid;
            else hi = mid - 1;
        }
        tile_of[tid] = lo;
    }
    __syncthreads();
    for (int w = tid; w < nt * kVSlots; w += kT) {
        const int j = w / kVSlots, e = w - j * kVSlots;
        if (e >= cnt[j]) continue;
        const int b = box[j], c = vl[kVSlots * (size_t)b + e];
        const int o = offset_index(ic[3 * (size_t)b] - ic[3 * (size_t)c], ic[3 * (size_t)b + 1] - ic[3 * (size_t)c + 1],
                                   ic[3 * (size_t)b + 2] - ic[3 * (size_t)c + 2]);
        const int row = src_row(o, par[j]);
        if (row >= 0) tab[row][j] = c;  // (always: V sources lie in the target's range)
    }
    __syncthreads();
    for (int e = tid; e < kSrcRows * kSrcTargets; e += kT) {
        const int k = e / kSrcTargets, j = e - k * kSrcTargets;
        if (j < nt) srcT[(size_t)k * np + r0 + j] = tab[k][j];
    }
    for (int k = tid; k < kNumOffsets; k += kT) {
        const int oi = kM2LSlots.offset_of[k];
        bool any = false;
        for (int j = 0; j < nt; ++j) {
            const int row = src_row(oi, par[j]);
            any |= row >= 0 && tab[row][j] >= 0;
            if (j + 1 == nt || tile_of[j + 1] != tile_of[j]) {  // end of a run in one tile
                if (any) atomicOr(&present[(size_t)tile_of[j] * kNumOffsets + k], 1);
                any = false;
            }
        }
    }
}

// Waves per M2L workgroup (W, a kernel shape the autotuner may choose): 16
// waves (4 per SIMD share the MFMA pipe and hide each other's LDS and gather
// latency; at p = 8 measured 8% faster than 8 waves of 32 targets) or 8. A
// wave owns one 16-target column block, except where one block's
// accumulators would not fit four waves per SIMD (FP64 above p = 8, split
// FP32): there 16-wave workgroups give each block two waves (m2l_row_split).
// The defaults per order and form are in default_m2l_waves.
constexpr int kM2LCb = 1;  // 16-target column blocks per wave

typedef float flt4 __attribute__((ext_vector_type(4)));

// Waves per 16-target column block of the FP64 / split-FP32 M2L workgroup:
// with 16 waves, two share each block (half the row blocks each) where one
// wave's accumulators would not fit four waves per SIMD (above p = 8, and
// the split-FP32 form with its FP64 fold totals).
constexpr int m2l_row_split(int p, bool mix, int waves) { return waves == 16 && (p > 8 || mix) ? 2 : 1; }
// Targets per M2L workgroup (tile) of order p, precision mode prec and waves
// per workgroup (the bf16 kernel has no row split).
inline int m2l_tile_targets(int p, int prec, int waves) {
    const int form = m2l_form(prec);
    const int rs = form == kPrecMixedBf16 ? 1 : m2l_row_split(p, form == kPrecMixed, waves);
    return waves / rs * 16 * kM2LCb;
}

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

__device__ __forceinline__ double a_value(double a) { return a; }
__device__ __forceinline__ double a_value(float2 a) { return (double)a.x + (double)a.y; }

// RSX (0: the default row split of the order and form) lets more waves share a
// column block when tiles are small (16 or 32 targets): each wave then runs a
// shorter chain per offset, and every row keeps its accumulation order.
template <int P, bool MIX, int W, int RSX = 0>
__global__ void __launch_bounds__(W * 64) k_m2l_mfma(int np, int ntiles, const int* __restrict__ tstart,
                                                           const int* __restrict__ order,
                                                           const int* __restrict__ ic,
                                                           const int* __restrict__ srcT,
                                                           const int* __restrict__ present_g,
                                                           const typename M2LTypes<MIX>::A* __restrict__ Asw,
                                                           const double* __restrict__ M, double* __restrict__ L) {
    using AT = typename M2LTypes<MIX>::A;
    using Acc = typename M2LTypes<MIX>::Acc;
    constexpr int S = nterms(P);
    constexpr int RB = m2l_mrb(S), TR = m2l_tail(S), NKC = m2l_mnkc(S), KC = m2l_mkc(S), CH = m2l_mchunk(S);
    constexpr int CB = kM2LCb;
    // RS waves share each 16-target column block, each taking RBH of the RB
    // row blocks (the tail rows go to the last of them): RS = 2 when a whole
    // column block's accumulators would not fit four waves per SIMD.
    constexpr int RS = RSX ? RSX : m2l_row_split(P, MIX, W);
    constexpr int RBH = (RB + RS - 1) / RS;
    constexpr int TILE = W / RS * 16 * CB;
    constexpr int NT = W * 64;
    constexpr int PER = (CH + NT - 1) / NT;
    // WHOLE: both buffers hold a whole operator (all NKC chunks), so the
    // workgroup synchronises once per offset; otherwise once per chunk.
    constexpr bool WHOLE = NKC > 1 && 2 * NKC * CH * (int)sizeof(AT) <= 150 * 1024;
    constexpr int BUFN = WHOLE ? NKC * CH : CH;
    __shared__ AT abuf[2][BUFN];
    __shared__ unsigned char present[kNumOffsets + 1];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int grp = wave / RS, rb0 = (wave % RS) * RBH;  // column block, first row block
    const bool tail = wave % RS == RS - 1;                  // this wave takes the tail rows
    const int j = lane & 15, kq = lane >> 4;
    const int tile = xcd_item();
    if (tile >= ntiles) return;
    const int base = tstart[tile], end = tstart[tile + 1];  // end - base <= TILE
    int rank[CB], par[CB];
    bool valid[CB];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
        rank[cb] = base + (grp * CB + cb) * 16 + j;
        valid[cb] = rank[cb] < end;
        par[cb] = valid[cb] ? ic_parity(ic, order[rank[cb]]) : 0;
    }
    Acc acc[CB][RBH];
    // MIX: each offset's products are summed in FP32 from zero and folded into
    // FP64 totals when the offset is done, so FP32 rounding never compounds
    // across the up to 189 offsets of a target.
    dbl4 accd[MIX ? CB : 1][MIX ? RBH : 1];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int rb = 0; rb < RBH; ++rb)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[cb][rb][r] = 0;
    if constexpr (MIX)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
#pragma unroll
            for (int rb = 0; rb < RBH; ++rb)
#pragma unroll
                for (int r = 0; r < 4; ++r) accd[cb][rb][r] = 0.0;
    (void)TILE;
    // tail rows 16 RB + tr: lane (j, kq) sums its k = kq (mod 4) terms
    double tacc[CB][TR > 0 ? TR : 1];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int tr = 0; tr < (TR > 0 ? TR : 1); ++tr) tacc[cb][tr] = 0.0;
    {
        for (int t = tid; t < kNumOffsets; t += NT) present[t] = present_g[(size_t)tile * kNumOffsets + t];
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
                const int row = o < kNumOffsets && valid[cb] ? src_row(kM2LSlots.offset_of[o], par[cb]) : -1;
                src[cb] = row >= 0 ? srcT[(size_t)row * np + rank[cb]] : -1;
                any |= src[cb] >= 0;
            }
            return __ballot(any) != 0ull;
        };
        int o = next_present(-1);
        int scur[CB], snext[CB];
        double bcur[CB][KC], bnext[CB][KC];
        bool has_cur = false, has_next = false;
        int onext = kNumOffsets;  // the offset after o
        if (o < kNumOffsets) {
            const AT* src = Asw + (size_t)o * NKC * CH;
            for (int e = tid; e < (WHOLE ? NKC * CH : CH); e += NT) abuf[0][e] = src[e];
            has_cur = load_src(o, scur);
            if (has_cur) load_b(scur, 0, bcur);
            onext = next_present(o);
            has_next = load_src(onext, snext);
        }
        __syncthreads();
        int buf = 0, c = 0;
        while (o < kNumOffsets) {
            int on = o, cn = c + 1;
            if (cn == NKC) {
                cn = 0;
                on = onext;
            }
            // operator prefetch: chunk c of the next offset (WHOLE), else the next item's chunk
            const int pf_o = WHOLE ? onext : on;
            const int pf_c = WHOLE ? c : cn;
            AT pre[PER];
            if (pf_o < kNumOffsets) {
                const AT* src = Asw + ((size_t)pf_o * NKC + pf_c) * CH;
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
                const AT* ab = abuf[buf] + (WHOLE ? c * CH : 0) + lane;
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
                        for (int rb = 0; rb < RBH; ++rb) {
                            if (RS > 1 && rb0 + rb >= RB) continue;  // (the last wave's short share)
                            const float2 a = ab[((rb0 + rb) * KC + kk) * 64];
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
                        for (int rb = 0; rb < RBH; ++rb) {
                            if (RS > 1 && rb0 + rb >= RB) continue;
                            const double a = ab[((rb0 + rb) * KC + kk) * 64];
#pragma unroll
                            for (int cb = 0; cb < CB; ++cb)
                                acc[cb][rb] =
                                    __builtin_amdgcn_mfma_f64_16x16x4f64(a, bcur[cb][kk], acc[cb][rb], 0, 0, 0);
                        }
                }
                if (TR > 0 && tail) {
                    const AT* tb = abuf[buf] + (WHOLE ? c * CH : 0) + RB * KC * 64 + kq;
#pragma unroll
                    for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                        for (int tr = 0; tr < TR; ++tr) {
                            const double a = a_value(tb[(tr * KC + kk) * 4]);
#pragma unroll
                            for (int cb = 0; cb < CB; ++cb) tacc[cb][tr] = fma(a, bcur[cb][kk], tacc[cb][tr]);
                        }
                }
                if constexpr (MIX)
                    if (cn == 0)  // offset o done: fold its FP32 sums into FP64
#pragma unroll
                        for (int cb = 0; cb < CB; ++cb)
#pragma unroll
                            for (int rb = 0; rb < RBH; ++rb)
#pragma unroll
                                for (int r = 0; r < 4; ++r) {
                                    accd[cb][rb][r] += (double)acc[cb][rb][r];
                                    acc[cb][rb][r] = 0.f;
                                }
            }
            if (pf_o < kNumOffsets) {
                AT* dst = abuf[buf ^ 1] + (WHOLE ? pf_c * CH : 0);
#pragma unroll
                for (int i = 0; i < PER; ++i) {
                    const int e = tid + NT * i;
                    if (e < CH) dst[e] = pre[i];
                }
            }
            if (!WHOLE || cn == 0) {
                __syncthreads();
                buf ^= 1;
            }
            if (cn == 0 && on < kNumOffsets) {
#pragma unroll
                for (int cb = 0; cb < CB; ++cb) scur[cb] = snext[cb];
                has_cur = has_next;
                onext = next_present(on);
                has_next = load_src(onext, snext);
            }
#pragma unroll
            for (int cb = 0; cb < CB; ++cb)
#pragma unroll
                for (int kk = 0; kk < KC; ++kk) bcur[cb][kk] = bnext[cb][kk];
            o = on;
            c = cn;
        }
    }
    if constexpr (TR > 0) {
        // (kq 0 + kq 1) + (kq 2 + kq 3), in every lane
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
#pragma unroll
            for (int tr = 0; tr < TR; ++tr) {
                tacc[cb][tr] += __shfl_xor(tacc[cb][tr], 16, 64);
                tacc[cb][tr] += __shfl_xor(tacc[cb][tr], 32, 64);
            }
    }
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
        if (!valid[cb]) continue;
        const int t = order[rank[cb]];
#pragma unroll
        for (int rb = 0; rb < RBH; ++rb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // D(i, j) of lane j + 16 kq, register r: i = kq + 4r (f64), 4kq + r (f32)
                const int R0 = 16 * (rb0 + rb);
                const int row = MIX ? R0 + 4 * kq + r : R0 + kq + 4 * r;
                if (rb0 + rb < RB && row < S) {
                    if constexpr (MIX) L[(size_t)t * S + row] = accd[cb][rb][r];
                    else L[(size_t)t * S + row] = acc[cb][rb][r];
                }
            }
        if constexpr (TR > 0)
            if (tail && kq == 0)
#pragma unroll
                for (int tr = 0; tr < TR; ++tr) L[(size_t)t * S + 16 * RB + tr] = tacc[cb][tr];
    }
}

// ---- M2L on bf16 matrix cores with three-piece split operands -------------------
// Each operand x is carried as bf16 pieces hi + mid + lo (x to ~2^-24
// relative); a 16 x 16 x 32 block takes the six products whose order is at or
// above 2^-24 (lo*hi, hi*lo, mid*mid, mid*hi, hi*mid, hi*hi) on
// v_mfma_f32_16x16x32_bf16 into an FP32 accumulator. Operand layout (16x16x32
// bf16): A[i][k] in lane i + 16 (k/8), element k%8; B[k][j] in lane j + 16 (k/8);
// D[i][j] in lane j + 16 (i/4), register i%4. Multipoles are split once per
// evaluation into Mbf[box][k block][piece][k/8 % 4] (8 bf16 per 16-byte unit).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned short bf16_rne(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ double bf16_to_double(unsigned short b) {
    return (double)__uint_as_float((unsigned)b << 16);
}

template <int P>
__global__ void __launch_bounds__(kT) k_split_bf16(int nb, const double* __restrict__ M, uint4* __restrict__ Mbf) {
    constexpr int S = nterms(P), KB = m2l_bf_kb(S);
    const int t = blockIdx.x * kT + threadIdx.x;
    if (t >= nb * KB * 4) return;
    const int box = t / (KB * 4), kb = (t / 4) % KB, kq = t % 4;
    unsigned short pc[kBfPieces][8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int k = 32 * kb + 8 * kq + e;
        double r = k < S ? M[(size_t)box * S + k] : 0.0;
#pragma unroll
        for (int q = 0; q < kBfPieces; ++q) {
            pc[q][e] = bf16_rne((float)r);
            r -= bf16_to_double(pc[q][e]);
        }
    }
#pragma unroll
    for (int q = 0; q < kBfPieces; ++q) {
        uint4 v;
        v.x = pc[q][0] | ((unsigned)pc[q][1] << 16);
        v.y = pc[q][2] | ((unsigned)pc[q][3] << 16);
        v.z = pc[q][4] | ((unsigned)pc[q][5] << 16);
        v.w = pc[q][6] | ((unsigned)pc[q][7] << 16);
        Mbf[(((size_t)box * KB + kb) * kBfPieces + q) * 4 + kq] = v;
    }
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// W waves of 16 targets (by default 16: measured 10% faster than 8 waves at p = 8)
template <int P, int W>
__global__ void __launch_bounds__(W * 64) k_m2l_bf16(int np, int ntiles, const int* __restrict__ tstart,
                                                           const int* __restrict__ order,
                                                           const int* __restrict__ ic,
                                                           const int* __restrict__ srcT,
                                                           const int* __restrict__ present_g,
                                                           const uint4* __restrict__ Abf,
                                                           const uint4* __restrict__ Mbf, double* __restrict__ L) {
    constexpr int S = nterms(P);
    constexpr int RB = m2l_rb(S), KB = m2l_bf_kb(S), NKC = m2l_bf_nkc(S), KC = m2l_bf_kc(S);
    constexpr int CHU = m2l_bf_chunk(S);
    constexpr int NT = W * 64;
    constexpr int PER = (CHU + NT - 1) / NT;
    __shared__ uint4 abuf[2][CHU];
    __shared__ unsigned char present[kNumOffsets + 1];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int j = lane & 15, kq = lane >> 4;
    const int tile = xcd_item();
    if (tile >= ntiles) return;
    const int base = tstart[tile], end = tstart[tile + 1];  // end - base <= 16 W
    const int rank = base + wave * 16 + j;
    const bool valid = rank < end;
    const int par = valid ? ic_parity(ic, order[rank]) : 0;
    // each offset's products are summed in FP32 from zero and folded into FP64
    // totals when the offset is done (FP32 rounding never compounds across offsets)
    flt4 acc[RB];
    dbl4 accd[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            acc[rb][r] = 0.f;
            accd[rb][r] = 0.0;
        }
    {
        for (int t = tid; t < kNumOffsets; t += NT) present[t] = present_g[(size_t)tile * kNumOffsets + t];
        if (tid == 0) present[kNumOffsets] = 1;  // sentinel
        __syncthreads();
        auto next_present = [&](int o) {
            ++o;
            while (!present[o]) ++o;
            return o;  // kNumOffsets when none left
        };
        auto load_b = [&](int src, int c, uint4 (&bv)[KC][kBfPieces]) {
            const uint4* ms = Mbf + (size_t)(src >= 0 ? src : 0) * KB * kBfPieces * 4;
#pragma unroll
            for (int kk = 0; kk < KC; ++kk) {
                const int kb = c * KC + kk;
#pragma unroll
                for (int q = 0; q < kBfPieces; ++q)
                    bv[kk][q] = (src >= 0 && kb < KB) ? ms[(kb * kBfPieces + q) * 4 + kq] : make_uint4(0, 0, 0, 0);
            }
        };
        auto load_src = [&](int o, int& src) {
            const int row = o < kNumOffsets && valid ? src_row(kM2LSlots.offset_of[o], par) : -1;
            src = row >= 0 ? srcT[(size_t)row * np + rank] : -1;
            return __ballot(src >= 0) != 0ull;
        };
        int o = next_present(-1);
        int scur = -1, snext = -1;
        uint4 bcur[KC][kBfPieces], bnext[KC][kBfPieces];
        bool has_cur = false, has_next = false;
        if (o < kNumOffsets) {
            const uint4* src = Abf + (size_t)o * NKC * CHU;
            for (int e = tid; e < CHU; e += NT) abuf[0][e] = src[e];
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
            uint4 pre[PER];
            if (on < kNumOffsets) {
                const uint4* src = Abf + ((size_t)on * NKC + cn) * CHU;
#pragma unroll
                for (int i = 0; i < PER; ++i) {
                    const int e = tid + NT * i;
                    pre[i] = (e < CHU) ? src[e] : make_uint4(0, 0, 0, 0);
                }
            }
            const bool nxt_has = (cn == 0) ? has_next : has_cur;
            if (on < kNumOffsets && nxt_has) load_b(cn == 0 ? snext : scur, cn, bnext);
            if (has_cur) {
                const uint4* ab = abuf[buf] + lane;
#pragma unroll
                for (int kk = 0; kk < KC; ++kk) {
                    const bf16x8 bh = as_bf16x8(bcur[kk][0]), bm = as_bf16x8(bcur[kk][1]),
                                 bl = as_bf16x8(bcur[kk][2]);
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb) {
                        const uint4* af = ab + ((rb * KC + kk) * kBfPieces) * 64;
                        const bf16x8 ah = as_bf16x8(af[0]), am = as_bf16x8(af[64]), al = as_bf16x8(af[128]);
                        flt4 a = acc[rb];
                        a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, a, 0, 0, 0);
                        a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, a, 0, 0, 0);
                        a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, a, 0, 0, 0);
                        a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, a, 0, 0, 0);
                        a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, a, 0, 0, 0);
                        acc[rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, a, 0, 0, 0);
                    }
                }
                if (cn == 0)  // offset o done
#pragma unroll
                    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            accd[rb][r] += (double)acc[rb][r];
                            acc[rb][r] = 0.f;
                        }
            }
            if (on < kNumOffsets) {
#pragma unroll
                for (int i = 0; i < PER; ++i) {
                    const int e = tid + NT * i;
                    if (e < CHU) abuf[buf ^ 1][e] = pre[i];
                }
            }
            __syncthreads();
            buf ^= 1;
            if (cn == 0 && on < kNumOffsets) {
                scur = snext;
                has_cur = has_next;
                has_next = load_src(next_present(on), snext);
            }
#pragma unroll
            for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                for (int q = 0; q < kBfPieces; ++q) bcur[kk][q] = bnext[kk][q];
            o = on;
            c = cn;
        }
    }
    if (!valid) return;
    const int t = order[rank];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * rb + 4 * kq + r;
            if (row < S) L[(size_t)t * S + row] = accd[rb][r];
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

// Degree of coefficient row r (rows n^2 .. (n+1)^2 - 1 hold degree n).
constexpr int degree_of(int r) {
    int n = 0;
    while ((n + 1) * (n + 1) <= r) ++n;
    return n;
}
// Whether the 16 x 4 operator block (row block rb, k step ks) can be nonzero:
// M2M takes sources of degree <= the target's, L2L of degree >= the target's.
template <int S, int MODE>
constexpr bool level_block_nonzero(int rb, int ks) {
    const int rlo = 16 * rb, clo = 4 * ks;
    if (rlo >= S || clo >= S) return false;
    const int rhi = rlo + 15 < S ? rlo + 15 : S - 1, chi = clo + 3 < S ? clo + 3 : S - 1;
    return MODE == 0 ? degree_of(clo) <= degree_of(rhi) : degree_of(chi) >= degree_of(rlo);
}

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
        // structurally zero blocks (a third of them at p = 8) are skipped
#pragma unroll
        for (int c = 0; c < NKC; ++c) {
            double bv[KC], av[RB * KC];
            const double* __restrict__ A = Tsw + ((size_t)o * NKC + c) * CH + lane;
#pragma unroll
            for (int e = 0; e < RB * KC; ++e)
                av[e] = level_block_nonzero<S, MODE>(e / KC, c * KC + e % KC) ? A[e * 64] : 0.0;
#pragma unroll
            for (int kk = 0; kk < KC; ++kk) {
                const int k = 4 * (c * KC + kk) + kq;
                bv[kk] = (src >= 0 && k < S) ? ms[k] : 0.0;
            }
#pragma unroll
            for (int kk = 0; kk < KC; ++kk)
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
                    if (level_block_nonzero<S, MODE>(rb, c * KC + kk))
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
// far = L2P + farw (the M2P field, when there are W lists); potential alike.
__global__ void __launch_bounds__(kT) k_far(TreeView t, const double* __restrict__ L, double* __restrict__ far,
                                            double* __restrict__ fpot, const double* __restrict__ farw,
                                            const double* __restrict__ fpotw) {
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
    double fx = gx * s2, fy = gy * s2, fz = gz * s2;
    if (farw) {
        fx += farw[3 * (size_t)i];
        fy += farw[3 * (size_t)i + 1];
        fz += farw[3 * (size_t)i + 2];
    }
    far[3 * (size_t)i] = fx;
    far[3 * (size_t)i + 1] = fy;
    far[3 * (size_t)i + 2] = fz;
    if (fpot) {
        double pt = local_pot<P>([&](int k) { return lb[k]; }, dx, dy, dz) * t.invh[b];
        if (fpotw) pt += fpotw[i];
        fpot[i] = pt;
    }
}

// L += Lx for the boxes with an X list (the P2L totals, after M2L).
__global__ void __launch_bounds__(kT) k_add_x(int nb, int S, const int* __restrict__ x_off,
                                              const double* __restrict__ Lx, double* __restrict__ L) {
    const long long g = (long long)blockIdx.x * kT + threadIdx.x;
    if (g >= (long long)nb * S) return;
    const int b = (int)(g / S);
    if (x_off[b + 1] > x_off[b]) L[g] += Lx[g];
}

// ---- P2P over U + final combine (AdaptiveFMM3D.hpp:96-106) ------------------
// One wave per target leaf. The source bodies of U(b) (contiguous ranges in
// sorted order) are streamed as one concatenated sequence in 64-body LDS
// tiles: the ranges of up to kStreamChunk U boxes and their running body
// offsets are staged in LDS once, and each lane finds the body of its tile
// slot by binary search. Every lane holds R targets (R = 2 above 32 bodies,
// else 1) of a lane group of T lanes (T = ceil(n_leaf / R), 8 <= T <= 64);
// the wave's G = floor(64/T) groups take the tile's sources g, g+G, ... and
// their partial sums are reduced in a fixed order (deterministic). The kernel
// is (|d|^2 + eps^2)^{-3/2} from v_rsq_f64 refined by one Newton step; the
// self pair is skipped by index, as in the reference (branch-free: its
// distance is replaced by 1 and its mass by 0), where its term is not already
// an exact zero (the potential, and eps = 0).
//
// v_rsq_f64 is accurate to 2^-24.2 relative on gfx950 (measured over 2^20
// arguments spanning 2^-20 .. 2^20); one Newton step brings that to 4.1e-15
// (a second one to 2.4e-16). The per-pair error after one step is far below
// the 1e-12 agreement asked of the FP64 forces, so the second step (4 FP64
// operations per pair) is not spent.
__device__ __forceinline__ double rsqrt_f64(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    return y * (1.5 - (0.5 * x) * y * y);
}

template <bool POT>
struct P2PAcc {
    double x = 0.0, y = 0.0, z = 0.0, p = 0.0;
};

// MIX: the pair difference stays FP64 (it carries the cancellation), then
// r^2, v_rsq_f32 and the kernel factor m / r^3 run in FP32; that factor times
// the FP64 difference is added to the target's FP64 accumulator pair by pair.
template <bool MIX, bool POT, int R, bool MASK>
__device__ __forceinline__ void p2p_tile(const double* sx, const double* sy, const double* sz, const double* sm,
                                         const int* sj, int kfirst, int cnt, int Gs, const double (&xi)[R][3],
                                         const int (&ii)[R], double eps2, P2PAcc<POT> (&a)[R]) {
    if constexpr (MIX) {
        const float e2 = (float)eps2;
#pragma unroll 1
        for (int k = kfirst; k < cnt; k += Gs) {
            const double xk = sx[k], yk = sy[k], zk = sz[k];
            const float mk = (float)sm[k];
            const int j = MASK ? sj[k] : 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double ddx = xk - xi[r][0], ddy = yk - xi[r][1], ddz = zk - xi[r][2];
                const float dx = (float)ddx, dy = (float)ddy, dz = (float)ddz;
                const bool self = MASK && j == ii[r];
                const float r2 = self ? 1.f : dx * dx + dy * dy + dz * dz + e2;
                const float m = self ? 0.f : mk;
                const float rin = __builtin_amdgcn_rsqf(r2);
                const double sc = (double)(m * (rin * rin * rin));
                a[r].x = fma(sc, ddx, a[r].x);
                a[r].y = fma(sc, ddy, a[r].y);
                a[r].z = fma(sc, ddz, a[r].z);
                if constexpr (POT) a[r].p += (double)(m * rin);
            }
        }
    } else {
#pragma unroll 1
        for (int k = kfirst; k < cnt; k += Gs) {
            const double xk = sx[k], yk = sy[k], zk = sz[k], mk = sm[k];
            const int j = MASK ? sj[k] : 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double dx = xk - xi[r][0], dy = yk - xi[r][1], dz = zk - xi[r][2];
                const bool self = MASK && j == ii[r];
                const double r2 = self ? 1.0 : dx * dx + dy * dy + dz * dz + eps2;
                const double m = self ? 0.0 : mk;
                const double rin = rsqrt_f64(r2);
                const double sc = m * (rin * rin * rin);
                a[r].x += sc * dx;
                a[r].y += sc * dy;
                a[r].z += sc * dz;
                if constexpr (POT) a[r].p += m * rin;
            }
        }
    }
}

// NEAR: write the near field (sorted order) to nr/npot for a later combine
// instead of combining with far here.
template <bool MIX, bool POT, int R, bool NEAR>
__device__ __forceinline__ void p2p_leaf(const TreeView& t, int b, int li, const double* __restrict__ far,
                                         double* __restrict__ acc, const double* __restrict__ fpot,
                                         double* __restrict__ phi, double* __restrict__ nr,
                                         double* __restrict__ npot, StreamLds& sl, double* sx, double* sy,
                                         double* sz, double* sm, int* sj, double (*red)[64], int min_group) {
    const int f = t.seg_first[b], nt = t.seg_count[b];
    const int lane = threadIdx.x;
    const int q = t.bsys[b];
    const double G = t.sys_G[q], eps2 = t.sys_eps2[q];
    // With eps^2 >= 1e-20 the self pair (and any coincident pair) has d = +0
    // and a finite kernel (also in FP32), so its acceleration term is exactly
    // +0 and adding it leaves every sum's bits unchanged: the index test and
    // masking are then skipped. They stay for the potential (the self term
    // m/eps is not zero) and for eps = 0.
    const bool mask = POT || !(eps2 >= 1e-20);
    const int ua = t.u_off[li], nu = t.u_off[li + 1] - ua;
    const int T = min(64, max(min_group, (nt + R - 1) / R));
    const int Gs = 64 / T;
    const int g = lane / T, tl = lane % T;
    const int kfirst = g < Gs ? g : 64;  // lanes past the last whole group idle
    for (int tb = 0; tb < nt; tb += R * T) {
        double xi[R][3];
        int ii[R];
        bool act[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = f + tb + r * T + tl;
            act[r] = g < Gs && tb + r * T + tl < nt;
            ii[r] = act[r] ? i : -2;
            xi[r][0] = act[r] ? t.x[i] : 0.0;
            xi[r][1] = act[r] ? t.y[i] : 0.0;
            xi[r][2] = act[r] ? t.z[i] : 0.0;
        }
        P2PAcc<POT> a[R];
        for (int e0 = 0; e0 < nu; e0 += kStreamChunk) {
            const int m = min(kStreamChunk, nu - e0);
            __syncthreads();  // previous readers of the staged ranges are done
            const int total = stage_chunk(sl, t.u_idx + ua, e0, m, t.seg_first, t.seg_count, lane);
            for (int q0 = 0; q0 < total; q0 += 64) {
                const int qq = q0 + lane;
                if (qq < total) {
                    const int j = stream_body(sl, m, qq);
                    sx[lane] = t.x[j];
                    sy[lane] = t.y[j];
                    sz[lane] = t.z[j];
                    sm[lane] = t.m[j];
                    sj[lane] = j;
                }
                __syncthreads();
                if (mask)
                    p2p_tile<MIX, POT, R, true>(sx, sy, sz, sm, sj, kfirst, min(64, total - q0), Gs, xi, ii, eps2, a);
                else
                    p2p_tile<MIX, POT, R, false>(sx, sy, sz, sm, sj, kfirst, min(64, total - q0), Gs, xi, ii, eps2, a);
                __syncthreads();
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (Gs > 1) {
                red[0][lane] = a[r].x;
                red[1][lane] = a[r].y;
                red[2][lane] = a[r].z;
                if constexpr (POT) red[3][lane] = a[r].p;
                __syncthreads();
                if (g == 0) {
                    for (int u = 1; u < Gs; ++u) {
                        a[r].x += red[0][u * T + tl];
                        a[r].y += red[1][u * T + tl];
                        a[r].z += red[2][u * T + tl];
                        if constexpr (POT) a[r].p += red[3][u * T + tl];
                    }
                }
                __syncthreads();
            }
            if (NEAR && act[r] && g == 0) {
                const int i = ii[r];
                nr[3 * (size_t)i] = a[r].x;
                nr[3 * (size_t)i + 1] = a[r].y;
                nr[3 * (size_t)i + 2] = a[r].z;
                if constexpr (POT) npot[i] = a[r].p;
            } else if (act[r] && g == 0) {
                const int i = ii[r];
                const int o = t.orig[i];
                acc[3 * (size_t)o] = G * (far[3 * (size_t)i] + a[r].x);
                acc[3 * (size_t)o + 1] = G * (far[3 * (size_t)i + 1] + a[r].y);
                acc[3 * (size_t)o + 2] = G * (far[3 * (size_t)i + 2] + a[r].z);
                if constexpr (POT) phi[o] = -G * (fpot[i] + a[r].p);
            }
        }
    }
}

// One wave per leaf. Leaves with more than pair_bodies bodies give each lane
// two targets; at least min_group lanes share one target's sources. Without
// `order` the grid covers the leaves in XCD chunks of tree order; with it the
// workgroups are persistent and take order[k] for k = blockIdx.x, +gridDim.x,
// ... (order lists the leaves by work bin, largest first, so the largest
// leaves start first and the small ones fill in behind them).
template <bool MIX, bool POT, bool NEAR>
__global__ void __launch_bounds__(64) k_p2p(TreeView t, const double* __restrict__ far, double* __restrict__ acc,
                                           const double* __restrict__ fpot, double* __restrict__ phi,
                                           double* __restrict__ nr, double* __restrict__ npot, int pair_bodies,
                                           int min_group, const int* __restrict__ order) {
    __shared__ double sx[64], sy[64], sz[64], sm[64];
    __shared__ int sj[64];
    __shared__ StreamLds sl;
    __shared__ double red[4][64];
    auto leaf = [&](int li) {
        const int b = t.leaf_ids[li];
        if (t.seg_count[b] > pair_bodies)
            p2p_leaf<MIX, POT, 2, NEAR>(t, b, li, far, acc, fpot, phi, nr, npot, sl, sx, sy, sz, sm, sj, red,
                                        min_group);
        else
            p2p_leaf<MIX, POT, 1, NEAR>(t, b, li, far, acc, fpot, phi, nr, npot, sl, sx, sy, sz, sm, sj, red,
                                        min_group);
    };
    if (order) {
        for (int k = blockIdx.x; k < t.nleaves; k += gridDim.x) {
            __syncthreads();  // the previous leaf's readers of the LDS tiles are done
            leaf(order[k]);
        }
    } else {
        const int li = xcd_chunk_item();
        if (li < t.nleaves) leaf(li);
    }
}

// Work bin of each leaf for the persistent P2P schedule: klo = 63 - bit
// length of (bodies x U-list bodies), so an ascending stable sort puts the
// largest bins first and keeps tree order inside a bin.
__global__ void __launch_bounds__(kT) k_p2p_work(TreeView t, uint64_t* __restrict__ khi, uint32_t* __restrict__ klo,
                                                 int* __restrict__ val) {
    const int li = blockIdx.x * kT + threadIdx.x;
    if (li >= t.nleaves) return;
    const int b = t.leaf_ids[li];
    long long src = 0;
    for (int e = t.u_off[li]; e < t.u_off[li + 1]; ++e) src += t.seg_count[t.u_idx[e]];
    const unsigned long long w = (unsigned long long)t.seg_count[b] * (unsigned long long)src;
    const int bits = w ? 64 - __builtin_clzll(w) : 0;
    khi[li] = 0;
    klo[li] = (uint32_t)(63 - bits);
    val[li] = li;
}

// acc = G (far + near) and phi = -G (fpot + npot) per body in input order:
// the expressions of the combining P2P, so that both paths give the same
// bits. One wave per leaf (G is per system, systems are per box).
template <bool POT>
__global__ void __launch_bounds__(64) k_combine(TreeView t, const double* __restrict__ far,
                                               const double* __restrict__ nr, const double* __restrict__ fpot,
                                               const double* __restrict__ npot, double* __restrict__ acc,
                                               double* __restrict__ phi) {
    const int li = xcd_chunk_item();
    if (li >= t.nleaves) return;
    const int b = t.leaf_ids[li];
    const int f = t.seg_first[b], nt = t.seg_count[b];
    const double G = t.sys_G[t.bsys[b]];
    for (int k = threadIdx.x; k < nt; k += 64) {
        const size_t i = (size_t)(f + k);
        const size_t o = (size_t)t.orig[i];
        acc[3 * o] = G * (far[3 * i] + nr[3 * i]);
        acc[3 * o + 1] = G * (far[3 * i + 1] + nr[3 * i + 1]);
        acc[3 * o + 2] = G * (far[3 * i + 2] + nr[3 * i + 2]);
        if constexpr (POT) phi[o] = -G * (fpot[i] + npot[i]);
    }
}

// Persistent P2P workgroups: enough one-wave workgroups for every SIMD of the
// 256 CUs to hold several.
constexpr int kP2PPersistentGrid = 256 * 4 * 4;

// order: the work-binned leaf order of the persistent schedule, or null.
template <bool NEAR>
void launch_p2p(hipStream_t s, const TreeView& t, bool mixed, const double* far, double* acc, const double* fpot,
                double* phi, double* nr, double* npot, const Shapes& sh, const int* order) {
    const dim3 pg(order ? std::max(1, std::min(t.nleaves, kP2PPersistentGrid)) : xcd_chunk_grid(t.nleaves)), pb(64);
    const int pb2 = sh.p2p_pair_bodies, mg = sh.p2p_min_group;
    if (mixed && phi)
        hipLaunchKernelGGL((k_p2p<true, true, NEAR>), pg, pb, 0, s, t, far, acc, fpot, phi, nr, npot, pb2, mg, order);
    else if (mixed)
        hipLaunchKernelGGL((k_p2p<true, false, NEAR>), pg, pb, 0, s, t, far, acc, fpot, phi, nr, npot, pb2, mg, order);
    else if (phi)
        hipLaunchKernelGGL((k_p2p<false, true, NEAR>), pg, pb, 0, s, t, far, acc, fpot, phi, nr, npot, pb2, mg, order);
    else
        hipLaunchKernelGGL((k_p2p<false, false, NEAR>), pg, pb, 0, s, t, far, acc, fpot, phi, nr, npot, pb2, mg,
                           order);
}

// M2P's leaf order: a workgroup's work is its rounds, ceil(n / T) * ceil(w / G)
// for a leaf of n bodies and w W entries with k_m2p's lane groups (T lanes,
// G = L / T groups), and the leaves are sorted by descending rounds (capped at
// 511), stable, so leaves of equal work keep tree order. On the 1M merger the
// rounds run up to 73 against a mean of 5.2 over the leaves with W, and the
// heaviest leaves, clustered in tree order, otherwise start late and finish
// last on their XCD (in a greedy model of the dispatch: a makespan of 108
// rounds in tree order, 76 heavy first, 70 ideal).
__global__ void __launch_bounds__(kT) k_m2p_work(TreeView t, int groups, uint64_t* __restrict__ khi,
                                                 uint32_t* __restrict__ klo, int* __restrict__ val) {
    const int li = blockIdx.x * kT + threadIdx.x;
    if (li >= t.nleaves) return;
    constexpr int L = kM2PWaves * 64;
    const int nw = t.w_off[li + 1] - t.w_off[li], nt = t.seg_count[t.leaf_ids[li]];
    int r = 0;
    if (nw > 0) {
        const int T = nt > L ? L : max(nt, L / groups), G = L / T;
        r = min((nt + T - 1) / T * ((nw + G - 1) / G), 511);
    }
    khi[li] = 0;
    klo[li] = (uint32_t)(511 - r);
    val[li] = li;
}

const int* m2p_order(Pool& pool, hipStream_t s, const TreeView& t, int groups, bool reserve_only) {
    const int nl = t.nleaves;
    SortBufs sb;
    for (int k = 0; k < 2; ++k) {
        sb.khi[k] = pool.get<uint64_t>(k ? "mw_khi1" : "mw_khi0", nl);
        sb.klo[k] = pool.get<uint32_t>(k ? "mw_klo1" : "mw_klo0", nl);
        sb.val[k] = pool.get<int>(k ? "mw_val1" : "mw_val0", nl);
    }
    sb.hist = pool.get<int>("mw_hist", 512 * (size_t)sort_blocks(nl));
    sb.scan_tmp = pool.get<int>("mw_scan", scan_tmp_ints(512 * sort_blocks(nl) + nl) + 16);
    if (reserve_only) return nullptr;
    hipLaunchKernelGGL(k_m2p_work, dim3(blocks(nl)), dim3(kT), 0, s, t, groups, sb.khi[0], sb.klo[0], sb.val[0]);
    return sb.val[radix_sort(sb, nl, 0, 1, s)];
}

// The persistent schedule's leaf order on stream s: work bins, one radix pass.
const int* p2p_order(Pool& pool, hipStream_t s, const TreeView& t, bool reserve_only) {
    const int nl = t.nleaves;
    SortBufs sb;
    for (int k = 0; k < 2; ++k) {
        sb.khi[k] = pool.get<uint64_t>(k ? "pw_khi1" : "pw_khi0", nl);
        sb.klo[k] = pool.get<uint32_t>(k ? "pw_klo1" : "pw_klo0", nl);
        sb.val[k] = pool.get<int>(k ? "pw_val1" : "pw_val0", nl);
    }
    sb.hist = pool.get<int>("pw_hist", 512 * (size_t)sort_blocks(nl));
    sb.scan_tmp = pool.get<int>("pw_scan", scan_tmp_ints(512 * sort_blocks(nl) + nl) + 16);
    if (reserve_only) return nullptr;
    hipLaunchKernelGGL(k_p2p_work, dim3(blocks(nl)), dim3(kT), 0, s, t, sb.khi[0], sb.klo[0], sb.val[0]);
    return sb.val[radix_sort(sb, nl, 0, 1, s)];
}

// M2L preparation: the transposed source table srcT[row * np + rank] (rank:
// the target's position in the M2L order, row: src_row of the offset for the
// target's parities; -1 where it has no source at that offset) and the
// per-tile offset presence table.
void m2l_prep(hipStream_t s, const TreeView& t, int np, int ntiles, const int* tstart, int* srcT, int* present) {
    const int nb = t.nb;
    hipLaunchKernelGGL(k_zero_int, dim3(blocks((long long)ntiles * kNumOffsets)), dim3(kT), 0, s, present,
                       (size_t)ntiles * kNumOffsets);
    hipLaunchKernelGGL(k_m2l_src, dim3(xcd_grid((nb + kSrcTargets - 1) / kSrcTargets)), dim3(kT), 0, s, nb, np,
                       t.m2l_order,
                       t.vl, t.nv, t.ic, srcT, tstart, ntiles, present);
}

// The M2L buffers of order P and precision prec, taken from the pool under
// the names run<P> uses.
template <int P>
void m2l_prep_for(Pool& pool, hipStream_t s, const TreeView& t, int prec, const Shapes& sh) {
    const int nb = t.nb, np = nb;
    (void)prec;
    const int TILE = m2l_tile_targets(P, prec, sh.m2l_waves);
    if (TILE != 256 && TILE != 128) throw std::runtime_error("M2L tile size without a plan");
    const int* tstart = TILE == 256 ? t.m2l_tiles256 : t.m2l_tiles128;
    const int ntiles = TILE == 256 ? t.m2l_ntiles256 : t.m2l_ntiles128;
    int* srcT = pool.get<int>("m2l_srcT", (size_t)kSrcRows * np);
    int* present = pool.get<int>("m2l_present", (size_t)ntiles * kNumOffsets);
    m2l_prep(s, t, np, ntiles, tstart, srcT, present);
}

// P2M into M, then M2M level by level (upward pass).
template <int P>
void upward(hipStream_t s, const TreeView& t, const TreeMeta& meta, const FmmTables& tab, double* M, hipEvent_t e0,
            hipEvent_t e1) {
    if (e0) AFMM_CHECK(hipEventRecord(e0, s));
    hipLaunchKernelGGL((k_accumulate<P, 0>), dim3(xcd_chunk_grid(t.nleaves)), dim3(64), 0, s, t, M);
    for (int l = t.max_level - 1; l >= 0; --l) {
        const int off = meta.nonleaf_off[l], cnt = meta.nonleaf_off[l + 1] - off;
        if (cnt > 0)
            hipLaunchKernelGGL((k_level_mfma<P, 0>), dim3((cnt + 15) / 16), dim3(kLevelWaves * 64), 0, s, cnt,
                               t.nonleaf_ids + off,
                               t.child_first, t.child_mask, t.parent, t.octant, tab.m2m_sw, M, M);
    }
    if (e1) AFMM_CHECK(hipEventRecord(e1, s));
}

// M2L on stream s for a tree whose M2L preparation is done: the bf16 split of
// the multipoles (bf16 form) and the M2L kernel of the order, precision form
// and workgroup shape; events e0/e1 (optional) just before and after the kernel.
template <int P>
void m2l_launch(Pool& pool, hipStream_t s, const TreeView& t, const FmmTables& tab, int prec, const Shapes& sh,
                hipEvent_t e0, hipEvent_t e1) {
    constexpr int S = nterms(P);
    const int form = m2l_form(prec);
    const int nb = t.nb, np = nb;
    double* M = pool.get<double>("Mexp", (size_t)nb * S);
    double* L = pool.get<double>("Lexp", (size_t)nb * S);
    const int TILE = m2l_tile_targets(P, prec, sh.m2l_waves);
    if (TILE != 256 && TILE != 128) throw std::runtime_error("M2L tile size without a plan");
    const int* tstart = TILE == 256 ? t.m2l_tiles256 : t.m2l_tiles128;
    const int ntiles = TILE == 256 ? t.m2l_ntiles256 : t.m2l_ntiles128;
    const int* srcT = pool.get<int>("m2l_srcT", (size_t)kSrcRows * np);
    const int* present = pool.get<int>("m2l_present", (size_t)ntiles * kNumOffsets);
    uint4* Mbf = form == kPrecMixedBf16 ? pool.get<uint4>("m2l_Mbf", (size_t)nb * m2l_bf_kb(S) * kBfPieces * 4)
                                        : nullptr;
    if (form == kPrecMixedBf16) {
        constexpr int KB = m2l_bf_kb(S);
        hipLaunchKernelGGL(k_split_bf16<P>, dim3(blocks((long long)nb * KB * 4)), dim3(kT), 0, s, nb, M, Mbf);
    }
    if (e0) AFMM_CHECK(hipEventRecord(e0, s));
    auto launch_m2l = [&](auto wtag) {
        constexpr int W = decltype(wtag)::value;
        const dim3 g(xcd_grid(ntiles)), b(W * 64);
        if (form == kPrecMixedBf16)
            hipLaunchKernelGGL((k_m2l_bf16<P, W>), g, b, 0, s, np, ntiles, tstart, t.m2l_order, t.ic, srcT, present,
                               tab.m2l_bf, Mbf, L);
        else if (form == kPrecMixed)
            hipLaunchKernelGGL((k_m2l_mfma<P, true, W>), g, b, 0, s, np, ntiles, tstart, t.m2l_order, t.ic, srcT,
                               present, tab.m2l_hl, M, L);
        else if constexpr (W == 16 && m2l_row_split(P, false, W) == 1) {
            // tiles of 16 or 32 targets (small trees): four or two waves per block
            if (t.m2l_t256 <= 16)
                hipLaunchKernelGGL((k_m2l_mfma<P, false, W, 4>), g, b, 0, s, np, ntiles, tstart, t.m2l_order, t.ic, srcT,
                                   present, tab.m2l_sw, M, L);
            else if (t.m2l_t256 <= 32)
                hipLaunchKernelGGL((k_m2l_mfma<P, false, W, 2>), g, b, 0, s, np, ntiles, tstart, t.m2l_order, t.ic, srcT,
                                   present, tab.m2l_sw, M, L);
            else
                hipLaunchKernelGGL((k_m2l_mfma<P, false, W>), g, b, 0, s, np, ntiles, tstart, t.m2l_order, t.ic, srcT,
                                   present, tab.m2l_sw, M, L);
        } else
            hipLaunchKernelGGL((k_m2l_mfma<P, false, W>), g, b, 0, s, np, ntiles, tstart, t.m2l_order, t.ic, srcT,
                               present, tab.m2l_sw, M, L);
    };
    if (sh.m2l_waves == 16) launch_m2l(std::integral_constant<int, 16>{});
    else launch_m2l(std::integral_constant<int, 8>{});
    if (e1) AFMM_CHECK(hipEventRecord(e1, s));
}

// P2L over X into Lx (then x_done) and M2P over W into farw (then w_done) on
// stream s; either buffer is null when its lists are empty. They read only the
// tree, the lists, the bodies and the multipoles, so they can run beside M2L.
template <int P>
void far_passes(Pool& pool, hipStream_t s, const TreeView& t, const double* M, double* Lx, double* farw,
                double* fpotw, hipEvent_t x_done = nullptr, hipEvent_t w_done = nullptr) {
    if (Lx) hipLaunchKernelGGL((k_accumulate<P, 2>), dim3(xcd_chunk_grid(t.nb)), dim3(64), 0, s, t, Lx);
    if (x_done) AFMM_CHECK(hipEventRecord(x_done, s));
    if (farw) {
        const int* ord = m2p_order(pool, s, t, m2p_groups<P>(), false);
        const dim3 mg(xcd_chunk_grid(t.nleaves)), mb(kM2PWaves * 64);
        if (fpotw) hipLaunchKernelGGL((k_m2p<P, true, false>), mg, mb, 0, s, t, M, farw, fpotw, ord);
        else hipLaunchKernelGGL((k_m2p<P, false, false>), mg, mb, 0, s, t, M, farw, fpotw, ord);
    }
    if (w_done) AFMM_CHECK(hipEventRecord(w_done, s));
}

// All device buffers are taken from the pool before the first launch, so that
// a call with reserve_only (no launches) sizes them and a following call under
// stream capture allocates nothing.
template <int P>
void run(Pool& pool, hipStream_t s, const TreeView& t, const TreeMeta& meta, const FmmTables& tab, int prec,
         const Shapes& sh, double* acc, double* phi, hipEvent_t* ev, bool reserve_only, const SideStream& side) {
    const bool mixed = p2p_fp32(prec);
    const int form = m2l_form(prec);
    constexpr int S = nterms(P);
    const int nb = t.nb;
    double* M = pool.get<double>("Mexp", (size_t)nb * S);
    double* L = pool.get<double>("Lexp", (size_t)nb * S);
    double* far = pool.get<double>("far", 3 * (size_t)t.n);
    double* fpot = phi ? pool.get<double>("farpot", t.n) : nullptr;
    // M2L targets in (octant, level) order, tiles cut inside octant groups
    const int TILE = m2l_tile_targets(P, prec, sh.m2l_waves);
    if (TILE != 256 && TILE != 128) throw std::runtime_error("M2L tile size without a plan");
    const int* tstart = TILE == 256 ? t.m2l_tiles256 : t.m2l_tiles128;
    const int ntiles = TILE == 256 ? t.m2l_ntiles256 : t.m2l_ntiles128;
    const int np = nb;
    int* srcT = pool.get<int>("m2l_srcT", (size_t)kSrcRows * np);
    int* present = pool.get<int>("m2l_present", (size_t)ntiles * kNumOffsets);
    uint4* Mbf = form == kPrecMixedBf16 ? pool.get<uint4>("m2l_Mbf", (size_t)nb * m2l_bf_kb(S) * kBfPieces * 4)
                                        : nullptr;
    (void)Mbf;  // (sized here for the capture; m2l_launch takes it)
    // with P2L and M2P beside M2L: their totals (added to L after M2L) and field
    // (added by L2P); in order they add into L and far directly (the same sums)
    const bool fb = side.far_beside;
    double* Lx = fb && t.n_x > 0 ? pool.get<double>("Lx", (size_t)nb * S) : nullptr;
    double* farw = fb && t.n_w > 0 ? pool.get<double>("farw", 3 * (size_t)t.n) : nullptr;
    double* fpotw = fb && t.n_w > 0 && phi ? pool.get<double>("farwpot", t.n) : nullptr;
    const bool split = side.s != nullptr;
    double* nr = split ? pool.get<double>("near", 3 * (size_t)t.n) : nullptr;
    double* npot = split && phi ? pool.get<double>("nearpot", t.n) : nullptr;
    if (sh.p2p_schedule) p2p_order(pool, s, t, true);  // its buffers too
    if (t.n_w > 0) m2p_order(pool, s, t, m2p_groups<P>(), true);
    if (reserve_only) return;
    // P2P needs only the tree and the U lists: on the second stream it runs
    // beside the far-field passes. It is forked before M2L in FP64 (it shares
    // the GPU with M2L's start) and after the bf16 M2L, whose short MFMA phase
    // it slowed more than it gained (measured both ways in both forms).
    auto fork_p2p = [&]() {
        AFMM_CHECK(hipEventRecord(side.fork, s));
        AFMM_CHECK(hipStreamWaitEvent(side.s, side.fork, 0));
        if (ev) AFMM_CHECK(hipEventRecord(ev[6], side.s));
        launch_p2p<true>(side.s, t, mixed, far, acc, fpot, phi, nr, npot, sh,
                         sh.p2p_schedule ? p2p_order(pool, side.s, t, false) : nullptr);
        if (ev) AFMM_CHECK(hipEventRecord(ev[7], side.s));
        AFMM_CHECK(hipEventRecord(side.join, side.s));
    };
    const bool p2p_after_m2l = form == kPrecMixedBf16;
    if (split && !p2p_after_m2l && !side.p2p_launched) fork_p2p();
    if (!side.upward_launched) upward<P>(s, t, meta, tab, M, ev ? ev[0] : nullptr, ev ? ev[1] : nullptr);
    if (!side.m2l_launched) {
        if (!side.prep_done) m2l_prep(s, t, np, ntiles, tstart, srcT, present);
        if (side.upward_launched && side.up_done) AFMM_CHECK(hipStreamWaitEvent(s, side.up_done, 0));  // M done
        m2l_launch<P>(pool, s, t, tab, prec, sh, ev ? ev[5] : nullptr, nullptr);
    }
    if (split && p2p_after_m2l && !side.p2p_launched) fork_p2p();
    if (ev) AFMM_CHECK(hipEventRecord(ev[2], s));
    // U, W and X lists built on the second stream (beside M2L) are needed from here
    if (side.lists_done) AFMM_CHECK(hipStreamWaitEvent(s, side.lists_done, 0));
    // P2L: beside M2L into Lx (added here), or here into L
    if (fb) {
        AFMM_CHECK(hipStreamWaitEvent(s, side.x_done, 0));
        if (Lx) hipLaunchKernelGGL(k_add_x, dim3(blocks((long long)nb * S)), dim3(kT), 0, s, nb, S, t.x_off, Lx, L);
    } else if (t.n_x > 0) {
        hipLaunchKernelGGL((k_accumulate<P, 1>), dim3(xcd_chunk_grid(nb)), dim3(64), 0, s, t, L);
    }
    for (int l = 1; l <= t.max_level; ++l) {
        const int off = meta.level_off[l], cnt = meta.level_off[l + 1] - off;
        hipLaunchKernelGGL((k_level_mfma<P, 1>), dim3((cnt + 15) / 16), dim3(kLevelWaves * 64), 0, s, cnt,
                           t.lo_order + off,
                           t.child_first, t.child_mask, t.parent, t.octant, tab.l2l_sw, L, L);
    }
    if (ev) AFMM_CHECK(hipEventRecord(ev[3], s));
    if (fb) AFMM_CHECK(hipStreamWaitEvent(s, side.w_done, 0));
    hipLaunchKernelGGL(k_far<P>, dim3(blocks(t.n)), dim3(kT), 0, s, t, L, far, fpot, farw, fpotw);
    if (!fb && t.n_w > 0) {  // M2P in order: added to far after L2P
        const int* ord = m2p_order(pool, s, t, m2p_groups<P>(), false);
        const dim3 mg(xcd_chunk_grid(t.nleaves)), mb(kM2PWaves * 64);
        if (fpot) hipLaunchKernelGGL((k_m2p<P, true, true>), mg, mb, 0, s, t, M, far, fpot, ord);
        else hipLaunchKernelGGL((k_m2p<P, false, true>), mg, mb, 0, s, t, M, far, fpot, ord);
    }
    if (split) {
        AFMM_CHECK(hipStreamWaitEvent(s, side.join, 0));
        const dim3 cg(xcd_chunk_grid(t.nleaves)), cb(64);
        if (phi) hipLaunchKernelGGL(k_combine<true>, cg, cb, 0, s, t, far, nr, fpot, npot, acc, phi);
        else hipLaunchKernelGGL(k_combine<false>, cg, cb, 0, s, t, far, nr, fpot, npot, acc, phi);
    } else {
        if (ev) AFMM_CHECK(hipEventRecord(ev[6], s));
        launch_p2p<false>(s, t, mixed, far, acc, fpot, phi, nullptr, nullptr, sh,
                          sh.p2p_schedule ? p2p_order(pool, s, t, false) : nullptr);
        if (ev) AFMM_CHECK(hipEventRecord(ev[7], s));
    }
    if (ev) AFMM_CHECK(hipEventRecord(ev[4], s));
    AFMM_CHECK(hipGetLastError());
}

}  // namespace

void run_upward(Pool& pool, hipStream_t s, const TreeView& tv, const TreeMeta& meta, int p, const FmmTables& tab,
                hipEvent_t e0, hipEvent_t e1) {
    double* M = pool.get<double>("Mexp", (size_t)tv.nb * nterms(p));
    switch (p) {
#define AFMM_CASE(PP) \
    case PP: upward<PP>(s, tv, meta, tab, M, e0, e1); return;
        AFMM_CASE(1) AFMM_CASE(2) AFMM_CASE(3) AFMM_CASE(4) AFMM_CASE(5) AFMM_CASE(6)
        AFMM_CASE(7) AFMM_CASE(8) AFMM_CASE(9) AFMM_CASE(10) AFMM_CASE(11) AFMM_CASE(12)
        AFMM_CASE(13) AFMM_CASE(14) AFMM_CASE(15) AFMM_CASE(16)
#undef AFMM_CASE
    }
    throw std::runtime_error("unsupported expansion order");
}

void run_m2l_prep(Pool& pool, hipStream_t s, const TreeView& tv, int p, int prec, const Shapes& sh) {
    switch (p) {
#define AFMM_CASE(PP) \
    case PP: m2l_prep_for<PP>(pool, s, tv, prec, sh); return;
        AFMM_CASE(1) AFMM_CASE(2) AFMM_CASE(3) AFMM_CASE(4) AFMM_CASE(5) AFMM_CASE(6)
        AFMM_CASE(7) AFMM_CASE(8) AFMM_CASE(9) AFMM_CASE(10) AFMM_CASE(11) AFMM_CASE(12)
        AFMM_CASE(13) AFMM_CASE(14) AFMM_CASE(15) AFMM_CASE(16)
#undef AFMM_CASE
    }
    throw std::runtime_error("unsupported expansion order");
}

void run_m2l(Pool& pool, hipStream_t s, const TreeView& tv, int p, const FmmTables& tab, int prec, const Shapes& sh,
             hipEvent_t e0, hipEvent_t e1) {
    switch (p) {
#define AFMM_CASE(PP) \
    case PP: m2l_launch<PP>(pool, s, tv, tab, prec, sh, e0, e1); return;
        AFMM_CASE(1) AFMM_CASE(2) AFMM_CASE(3) AFMM_CASE(4) AFMM_CASE(5) AFMM_CASE(6)
        AFMM_CASE(7) AFMM_CASE(8) AFMM_CASE(9) AFMM_CASE(10) AFMM_CASE(11) AFMM_CASE(12)
        AFMM_CASE(13) AFMM_CASE(14) AFMM_CASE(15) AFMM_CASE(16)
#undef AFMM_CASE
    }
    throw std::runtime_error("unsupported expansion order");
}

void run_p2p_near(Pool& pool, hipStream_t s, const TreeView& tv, int prec, const Shapes& sh, bool pot, hipEvent_t e0,
                  hipEvent_t e1) {
    double* nr = pool.get<double>("near", 3 * (size_t)tv.n);
    double* npot = pot ? pool.get<double>("nearpot", tv.n) : nullptr;
    const int* order = sh.p2p_schedule ? p2p_order(pool, s, tv, false) : nullptr;
    if (e0) AFMM_CHECK(hipEventRecord(e0, s));
    // (the near-field kernel reads neither far nor acc; its potential flag is
    // the npot buffer)
    launch_p2p<true>(s, tv, p2p_fp32(prec), nullptr, nullptr, nullptr, npot, nr, npot, sh, order);
    if (e1) AFMM_CHECK(hipEventRecord(e1, s));
    AFMM_CHECK(hipGetLastError());
}

void run_far_passes(Pool& pool, hipStream_t s, const TreeView& tv, int p, bool pot, hipEvent_t x_done,
                    hipEvent_t w_done) {
    const int S = nterms(p);
    const double* M = pool.get<double>("Mexp", (size_t)tv.nb * S);
    double* Lx = tv.n_x > 0 ? pool.get<double>("Lx", (size_t)tv.nb * S) : nullptr;
    double* farw = tv.n_w > 0 ? pool.get<double>("farw", 3 * (size_t)tv.n) : nullptr;
    double* fpotw = tv.n_w > 0 && pot ? pool.get<double>("farwpot", tv.n) : nullptr;
    switch (p) {
#define AFMM_CASE(PP) \
    case PP: far_passes<PP>(pool, s, tv, M, Lx, farw, fpotw, x_done, w_done); return;
        AFMM_CASE(1) AFMM_CASE(2) AFMM_CASE(3) AFMM_CASE(4) AFMM_CASE(5) AFMM_CASE(6)
        AFMM_CASE(7) AFMM_CASE(8) AFMM_CASE(9) AFMM_CASE(10) AFMM_CASE(11) AFMM_CASE(12)
        AFMM_CASE(13) AFMM_CASE(14) AFMM_CASE(15) AFMM_CASE(16)
#undef AFMM_CASE
    }
    throw std::runtime_error("unsupported expansion order");
}

void run_fmm(Pool& pool, hipStream_t s, const TreeView& tv, const TreeMeta& meta, int p, const FmmTables& tab,
             int prec, const Shapes& sh, double* d_acc, double* d_phi, hipEvent_t* ev, bool reserve_only,
             SideStream side) {
    if (m2l_form(prec) == kPrecMixed && !tab.m2l_hl) throw std::runtime_error("split FP32 M2L operators not loaded");
    if (m2l_form(prec) == kPrecMixedBf16 && !tab.m2l_bf) throw std::runtime_error("bf16 M2L operators not loaded");
    switch (p) {
#define AFMM_CASE(PP) \
    case PP: run<PP>(pool, s, tv, meta, tab, prec, sh, d_acc, d_phi, ev, reserve_only, side); return;
        AFMM_CASE(1) AFMM_CASE(2) AFMM_CASE(3) AFMM_CASE(4) AFMM_CASE(5) AFMM_CASE(6)
        AFMM_CASE(7) AFMM_CASE(8) AFMM_CASE(9) AFMM_CASE(10) AFMM_CASE(11) AFMM_CASE(12)
        AFMM_CASE(13) AFMM_CASE(14) AFMM_CASE(15) AFMM_CASE(16)
#undef AFMM_CASE
    }
    throw std::runtime_error("unsupported expansion order");
}

}  // namespace afmm
