// ============================================================================
// tree.hip -- device build of the adaptive octree and its CGR interaction
// lists, bit-identical to AdaptiveOctree::build / buildLists
// (AdaptiveOctree.hpp:65-189).
//
// Why it matches the recursive reference exactly:
//  * Box centres are produced by the same IEEE operations along each path
//    (c_child = c_parent +- 0.5*half, AdaptiveOctree.hpp:121-127), so a body
//    that replays its own path from the root sees the very centres the
//    reference compares it with (x >= cx, :109-111). The path is kept as a
//    72-bit key of octant digits, level 1 most significant.
//  * The reference's stable counting sort per split leaves the bodies of every
//    box sorted by octant path, ties by input index. Pass 1 sorts by the key
//    (stable LSD radix); the topology then follows from the sorted keys
//    level by level (a box splits iff count > cap, 2*half > eps floor and
//    level < 24, :102-104). Inside a leaf the reference keeps input order, so
//    pass 2 re-sorts stably by leaf start.
//  * Reference box ids come from its allocation order: children are numbered
//    as a block when their parent is visited in DFS pre-order. With
//    pre(b) = DFS pre-order rank, id(child r of P) = 1 + sum_{pre(Q) < pre(P)}
//    nchild(Q) + r.
//  * Lists use the reference's adjacency test on those centres
//    (|dc| <= half_a + half_b + tol per axis, :58-62), and emit in the
//    reference orders: colleagues and V ascending id (= Morton order in a
//    level), U and W in DFS order, X ascending leaf id.
// Internally boxes are numbered breadth-first (level-major, Morton inside a
// level): children of a box are contiguous and each level is one range.
// ============================================================================
#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <mutex>

#include "afmm_internal.h"
#include "block_scan.h"
#include "afmm_math.h"

#pragma clang fp contract(off)

namespace afmm {
namespace {

constexpr int kT = 256;
inline int blocks(long long n, int t = kT) { return (int)std::max<long long>(1, (n + t - 1) / t); }

// ---- bounding cube ----------------------------------------------------------
// Sequential std::min/std::max keep the first of equal values (only a strict
// comparison replaces), which matters for the sign of a zero coordinate.
struct Ext {
    double v;
    int i;
};
__device__ __forceinline__ void take_lo(Ext& a, Ext b) {
    if (b.v < a.v || (b.v == a.v && b.i < a.i)) a = b;
}
__device__ __forceinline__ void take_hi(Ext& a, Ext b) {
    if (b.v > a.v || (b.v == a.v && b.i < a.i)) a = b;
}

constexpr int kBBoxChunk = 2048;  // about 500 blocks at 1M bodies

// One block per chunk [cbeg[c], cend[c]) of one system's bodies.
__global__ void __launch_bounds__(kT) k_bbox_partial(const double* pos, const double* mass,
                                                    const int* __restrict__ cbeg, const int* __restrict__ cend,
                                                    double* pv, int* pi, int* pbad) {
    Ext e[6];
    for (int a = 0; a < 3; ++a) {
        e[a] = Ext{1e300, -1};
        e[3 + a] = Ext{-1e300, -1};
    }
    int bad = 0;
    const int i0 = cbeg[blockIdx.x], i1 = cend[blockIdx.x];
    // all loads of the chunk first (chunks hold at most kBBoxChunk bodies),
    // then the comparisons: (value, index) order makes the result independent
    // of the order in which they are taken
    constexpr int kPer = kBBoxChunk / kT;
    double v[kPer][3], mv[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = i0 + threadIdx.x + k * kT;
        const bool in = i < i1;
#pragma unroll
        for (int a = 0; a < 3; ++a) v[k][a] = in ? pos[3 * (size_t)i + a] : 0.0;
        mv[k] = in ? mass[i] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int i = i0 + threadIdx.x + k * kT;
        if (i >= i1) continue;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            bad += !isfinite(v[k][a]);
            take_lo(e[a], Ext{v[k][a], i});
            take_hi(e[3 + a], Ext{v[k][a], i});
        }
        bad += !isfinite(mv[k]);
    }
    __shared__ double sv[6][kT];
    __shared__ int si[6][kT];
    __shared__ int sb[kT];
    for (int k = 0; k < 6; ++k) sv[k][threadIdx.x] = e[k].v, si[k][threadIdx.x] = e[k].i;
    sb[threadIdx.x] = bad;
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            const int o = threadIdx.x + off;
            for (int k = 0; k < 6; ++k) {
                Ext a{sv[k][threadIdx.x], si[k][threadIdx.x]}, b{sv[k][o], si[k][o]};
                if (k < 3) take_lo(a, b);
                else take_hi(a, b);
                sv[k][threadIdx.x] = a.v, si[k][threadIdx.x] = a.i;
            }
            sb[threadIdx.x] += sb[o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 6; ++k) pv[k * gridDim.x + blockIdx.x] = sv[k][0], pi[k * gridDim.x + blockIdx.x] = si[k][0];
        pbad[blockIdx.x] = sb[0];
    }
}

// AdaptiveOctree::build (:71-81) of system blockIdx.x from its chunks'
// partials [cfirst[q], cfirst[q+1]); also the key depth.
__global__ void __launch_bounds__(kT) k_bbox_final(const double* pv, const int* pi, const int* pbad, int nparts,
                                                  const int* __restrict__ cfirst,
                                                  const double* __restrict__ eps_floor, RootInfo* roots) {
    Ext e[6];
    for (int a = 0; a < 3; ++a) e[a] = Ext{1e300, -1}, e[3 + a] = Ext{-1e300, -1};
    int bad = 0;
    const int q = blockIdx.x;
    for (int b = cfirst[q] + threadIdx.x; b < cfirst[q + 1]; b += kT) {
        for (int k = 0; k < 6; ++k) {
            const Ext x{pv[k * nparts + b], pi[k * nparts + b]};
            if (k < 3) take_lo(e[k], x);
            else take_hi(e[k], x);
        }
        bad += pbad[b];
    }
    __shared__ double sv[6][kT];
    __shared__ int si[6][kT];
    __shared__ int sb[kT];
    for (int k = 0; k < 6; ++k) sv[k][threadIdx.x] = e[k].v, si[k][threadIdx.x] = e[k].i;
    sb[threadIdx.x] = bad;
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            const int o = threadIdx.x + off;
            for (int k = 0; k < 6; ++k) {
                Ext x{sv[k][threadIdx.x], si[k][threadIdx.x]}, y{sv[k][o], si[k][o]};
                if (k < 3) take_lo(x, y);
                else take_hi(x, y);
                sv[k][threadIdx.x] = x.v, si[k][threadIdx.x] = x.i;
            }
            sb[threadIdx.x] += sb[o];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    RootInfo r;
    for (int a = 0; a < 3; ++a) r.c[a] = 0.5 * (sv[a][0] + sv[3 + a][0]);
    double ext = sv[3][0] - sv[0][0];
    if (ext < sv[4][0] - sv[1][0]) ext = sv[4][0] - sv[1][0];
    if (ext < sv[5][0] - sv[2][0]) ext = sv[5][0] - sv[2][0];
    double half = 0.5 * ext;
    if (!(half > 0)) half = 1.0;
    half *= 1.0 + 1e-9;
    r.half = half;
    r.tol = 1e-9 * half;
    // Levels 0..L-1 may split (2*half_l > eps floor, level < 24): keys carry L digits.
    int L = 0;
    double h = half;
    while (L < kMaxDepth && 2.0 * h > eps_floor[q]) {
        ++L;
        h = 0.5 * h;
    }
    r.key_levels = L;
    r.nonfinite = sb[0];
    roots[q] = r;
}

// ---- path keys ----------------------------------------------------------------
// digit of level l (1..24): levels 1..21 in khi bits [63-3l, 66-3l), 22..24 in klo.
__device__ __forceinline__ int level_digit(uint64_t khi, uint32_t klo, int l) {
    return l <= 21 ? int((khi >> (63 - 3 * l)) & 7u) : int((klo >> (3 * (24 - l))) & 7u);
}

// System of body i: the q with sys_off[q] <= i < sys_off[q+1].
__device__ __forceinline__ int system_of(int i, const int* sys_off, int nsys) {
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sys_off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(kT) k_keys(const double* pos, int n, int nsys, const int* __restrict__ sys_off,
                                            const RootInfo* roots, uint64_t* khi, uint32_t* klo, int* val) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const int q = system_of(i, sys_off, nsys);
    const RootInfo r = roots[q];
    const double x = pos[3 * (size_t)i], y = pos[3 * (size_t)i + 1], z = pos[3 * (size_t)i + 2];
    double cx = r.c[0], cy = r.c[1], cz = r.c[2], h = r.half;
    uint64_t hi = 0;
    uint32_t lo = 0;
    for (int l = 1; l <= r.key_levels; ++l) {
        const int o = (x >= cx ? 1 : 0) | (y >= cy ? 2 : 0) | (z >= cz ? 4 : 0);
        h = 0.5 * h;
        cx = cx + ((o & 1) ? h : -h);
        cy = cy + ((o & 2) ? h : -h);
        cz = cz + ((o & 4) ? h : -h);
        if (l <= 21) hi |= uint64_t(o) << (63 - 3 * l);
        else lo |= uint32_t(o) << (3 * (24 - l));
    }
    khi[i] = hi;
    klo[i] = lo | (uint32_t(q) << 9);
    val[i] = i;
}

// ---- topology, one level at a time --------------------------------------------
// For each box of level l: split decision and child ranges on the level-(l+1)
// digit inside its (sorted) segment. Eight lanes per box (one wave holds eight
// boxes): lane o finds the first body with digit >= o by its own binary search
// over the segment, and the box's child count is gathered with shuffles.
__global__ void __launch_bounds__(kT) k_split(int off, int cnt, int l, int cap, const double* eps_floor,
                                             const RootInfo* roots, const int* bsys, const int* seg_first,
                                             const int* seg_count, const uint64_t* khi, const uint32_t* klo,
                                             int* bnd, int* nch) {
    const int gt = blockIdx.x * kT + threadIdx.x;
    const int i = gt >> 3, o = gt & 7;
    const bool live = i < cnt;
    int f = 0, c = 0;
    bool split = false;
    if (live) {
        const int b = off + i;
        f = seg_first[b];
        c = seg_count[b];
        const int q = bsys[b];
        const double half = ldexp(roots[q].half, -l);
        split = c > cap && !(2.0 * half <= eps_floor[q]) && l < kMaxDepth;
    }
    int lo = f;
    if (split && o > 0) {
        int hi = f + c;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (level_digit(khi[mid], klo[mid], l + 1) < o) lo = mid + 1;
            else hi = mid;
        }
    }
    // bound of octant o + 1 (lane o + 1 of the group; the last lane uses the segment end)
    const int up = __shfl_down(lo, 1, 8);
    const int end = (o == 7) ? f + c : up;
    int nonempty = (split && end > lo) ? 1 : 0;
#pragma unroll
    for (int m = 1; m < 8; m <<= 1) nonempty += __shfl_xor(nonempty, m, 8);
    if (!live) return;
    int* bd = bnd + 9 * (size_t)i;
    if (split) {
        bd[o] = lo;
        if (o == 7) bd[8] = f + c;
    }
    if (o == 0) nch[i] = split ? nonempty : 0;
}

__global__ void __launch_bounds__(kT) k_emit(int off, int cnt, int l, int next_off, const int* bnd,
                                            const int* nch, const int* cbase, int* seg_first,
                                            int* seg_count, int* level, int* parent, int* octant,
                                            int* child_first, int* nchild, unsigned char* mask, int* bsys) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= cnt) return;
    const int b = off + i;
    const int k = nch[i];
    nchild[b] = k;
    if (k == 0) {
        child_first[b] = -1;
        mask[b] = 0;
        return;
    }
    const int* bd = bnd + 9 * (size_t)i;
    const int first = next_off + cbase[i];
    child_first[b] = first;
    unsigned char mk = 0;
    int r = 0;
    for (int o = 0; o < 8; ++o) {
        if (bd[o + 1] <= bd[o]) continue;
        const int c = first + r++;
        mk |= (unsigned char)(1u << o);
        seg_first[c] = bd[o];
        seg_count[c] = bd[o + 1] - bd[o];
        level[c] = l + 1;
        parent[c] = b;
        octant[c] = o;
        bsys[c] = bsys[b];
    }
    mask[b] = mk;
}

// ---- topology of all levels at once -----------------------------------------------
// The run of sorted bodies that shares body i's first l digits holds more than
// cap bodies iff some window of cap + 1 consecutive bodies around i shares
// them, i.e. iff M(i) = max over s in [i - cap, i] of common(s, s + cap) >= l.
// A box splits iff it holds more than cap bodies and its level is below the
// system's key_levels (= the eps-floor and depth rule of the split), so the
// leaf holding i is at level d(i) = min(key_levels, M(i) + 1), and i is the
// first body of its boxes at levels a(i) < l <= d(i), a(i) = common(i - 1, i)
// (-1 at the first body of a system). Boxes of one level are numbered in body
// order, which is the level-by-level order of the split (children by parent,
// then octant): a box's id is its level's offset plus the number of bodies
// before i that start a box of that level. The box of level l holding body
// e - 1 ends at e iff a(e) < l; it is the last box of that level started
// before e.
constexpr int kTopoTile = 256;                     // bodies per block
constexpr int kTopoWinMax = 4096 - kTopoTile;      // largest leaf capacity of this path
constexpr int kTopoItems = (kTopoTile + kTopoWinMax) / kTopoTile;
static_assert(kTopoTile == 4 * kWave, "four waves per tile");
// Counts the passes gather for the host with the level offsets: leaves per
// level [0, kMaxDepth], boxes per octant, the largest leaf's body count.
// Also whether some body's leaf lies below the levels the topology passes took
// (nlev, the digits the sort ordered): then the build is redone with all digits.
constexpr int kTopoLeafCnt = 0, kTopoOctCnt = kMaxDepth + 1, kTopoMaxLeaf = kTopoOctCnt + 8;
constexpr int kTopoDeep = kTopoMaxLeaf + 1, kTopoExtra = kTopoDeep + 1;

// Levels (digits) two keys of one system share; -1 for keys of two systems.
// Level l's digit sits at khi bits [63-3l, 66-3l) (bit 63 is always 0) for
// l <= 21, at klo bits [3(24-l), 3(25-l)) above that; klo >> 9 is the system.
__device__ __forceinline__ int common_levels(uint64_t ha, uint32_t la, uint64_t hb, uint32_t lb) {
    if ((la ^ lb) >> 9) return -1;
    const uint64_t x = ha ^ hb;
    if (x) return (__clzll((long long)x) - 1) / 3;
    const uint32_t y = (la ^ lb) & 0x1FFu;
    if (y) return 21 + (__clz((int)y) - 23) / 3;
    return kMaxDepth;
}

// d(i) and a(i) + 1 per body, the number of boxes it starts (the reference-id
// pass reads these), and per level the bodies of this tile that start a box.
__global__ void __launch_bounds__(kTopoTile) k_topo_count(int n, int cap, int nlev, int ntiles,
                                                          const RootInfo* __restrict__ roots,
                                                          const uint64_t* __restrict__ khi,
                                                          const uint32_t* __restrict__ klo, int* __restrict__ dlev,
                                                          int* __restrict__ minlev, int* __restrict__ nstart,
                                                          int* __restrict__ tcnt, int* __restrict__ ovf,
                                                          int* __restrict__ part) {
    __shared__ int m[kTopoTile + kTopoWinMax];
    __shared__ int wc[kMaxDepth + 1][kTopoTile / kWave];
    const int tid = threadIdx.x, base = blockIdx.x * kTopoTile;
    const int lw = kTopoTile + cap;  // windows s = base - cap + x, x < lw
    for (int x = tid; x < lw; x += kTopoTile) {
        const int s = base - cap + x;
        int w = -1;
        if (s >= 0 && s + cap < n) w = common_levels(khi[s], klo[s], khi[s + cap], klo[s + cap]);
        m[x] = w;
    }
    // maxima over [x, x + P) by doubling, P the largest power of two <= cap + 1;
    // a window starting past lw - P is partial and never read below
    int P = 1;
    while (2 * P <= cap + 1) {
        __syncthreads();
        int v[kTopoItems];
#pragma unroll
        for (int k = 0; k < kTopoItems; ++k) {
            const int x = tid + k * kTopoTile;
            v[k] = x < lw ? max(m[x], x + P < lw ? m[x + P] : -1) : -1;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kTopoItems; ++k) {
            const int x = tid + k * kTopoTile;
            if (x < lw) m[x] = v[k];
        }
        P *= 2;
    }
    __syncthreads();
    const int i = base + tid;
    int a = kMaxDepth, d = -1;  // a body past the end starts nothing
    bool deep = false;
    if (i < n) {
        const int M = max(m[tid], m[tid + cap + 1 - P]);  // windows [tid, tid + cap]
        const uint64_t h = khi[i];
        const uint32_t lo = klo[i];
        // (nlev below the system's own depth, a guess, stops every body there:
        // such a build is only checked and then redone, but it stays consistent)
        const int full = min(roots[lo >> 9].key_levels, M + 1);
        deep = full > nlev;
        d = min(full, nlev);
        a = i > 0 ? common_levels(khi[i - 1], klo[i - 1], h, lo) : -1;
        dlev[i] = d;
        minlev[i] = a + 1;
        nstart[i] = max(0, d - a);
    }
    const int lane = tid & (kWave - 1), wv = tid / kWave;
    for (int l = 0; l <= nlev; ++l) {
        const uint64_t b = __ballot(a < l && l <= d);
        if (lane == 0) wc[l][wv] = __popcll(b);
    }
    __syncthreads();
    if (tid <= nlev) tcnt[(size_t)tid * ntiles + blockIdx.x] = wc[tid][0] + wc[tid][1] + wc[tid][2] + wc[tid][3];
    if (blockIdx.x == 0 && tid == 0) *ovf = 0;
    const int anydeep = __syncthreads_or(deep ? 1 : 0);
    if (tid == 0) part[(size_t)kTopoDeep * ntiles + blockIdx.x] = anydeep;
}

// Per level (one block each): exclusive scan of the tile counts in place, and
// the level's total.
__global__ void __launch_bounds__(kScanThreads) k_topo_scan(int ntiles, int* __restrict__ tcnt,
                                                            int* __restrict__ ltot) {
    __shared__ int lds[kScanThreads / kWave];
    int* row = tcnt + (size_t)blockIdx.x * ntiles;
    int carry = 0;
    for (int c0 = 0; c0 < ntiles; c0 += kScanTile) {
        int v[kScanItems], sum = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int t = c0 + threadIdx.x * kScanItems + k;
            v[k] = t < ntiles ? row[t] : 0;
            sum += v[k];
        }
        int total;
        int ex = block_exclusive_scan(sum, lds, total) + carry;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int t = c0 + threadIdx.x * kScanItems + k;
            if (t < ntiles) row[t] = ex;
            ex += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) ltot[blockIdx.x] = carry;
}

// Level offsets (off[0..nlev+1]) and per-level wave counts of the tile's
// starts: the shared part of the two passes below. Returns through `off` and
// `wc` (LDS) after a barrier.
__device__ __forceinline__ void topo_tile_setup(int nlev, const int* __restrict__ ltot, int a, int d, int* off,
                                                int (*wc)[kTopoTile / kWave]) {
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    if (tid == 0) {
        int acc = 0;
        for (int l = 0; l <= nlev; ++l) {
            off[l] = acc;
            acc += ltot[l];
        }
        off[nlev + 1] = acc;
    }
    for (int l = 0; l <= nlev; ++l) {
        const uint64_t b = __ballot(a < l && l <= d);
        if (lane == 0) wc[l][wv] = __popcll(b);
    }
    __syncthreads();
}

// Number of bodies before this thread's body that start a box of level l
// (this body's box id there if it starts one), given the level's ballot.
__device__ __forceinline__ int topo_rank(int l, uint64_t b, const int* off, const int (*wc)[kTopoTile / kWave],
                                         const int* __restrict__ tpre, int ntiles) {
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int r = off[l] + tpre[(size_t)l * ntiles + blockIdx.x] + __popcll(b & lt);
    for (int w = 0; w < wv; ++w) r += wc[l][w];
    return r;
}

// Every box: first body, level, octant, system, parent, first child; leaves
// get no children. Box ids past `capacity` set *ovf and are not written.
__global__ void __launch_bounds__(kTopoTile) k_topo_emit(int n, int nlev, int ntiles, int capacity,
                                                         const uint64_t* __restrict__ khi,
                                                         const uint32_t* __restrict__ klo,
                                                         const int* __restrict__ dlev, const int* __restrict__ minlev,
                                                         const int* __restrict__ tpre, const int* __restrict__ ltot,
                                                         int* __restrict__ lev, int* ovf, int* seg_first, int* level,
                                                         int* parent, int* octant, int* child_first, int* nchild,
                                                         unsigned char* mask, int* bsys, int* __restrict__ part) {
    __shared__ int off[kMaxDepth + 2];
    __shared__ int wc[kMaxDepth + 1][kTopoTile / kWave];
    __shared__ int cnt[kTopoMaxLeaf];  // leaves per level, then boxes per octant
    const int e = blockIdx.x * kTopoTile + threadIdx.x;
    if (threadIdx.x < kTopoMaxLeaf) cnt[threadIdx.x] = 0;  // (read after the barriers below)
    int a = kMaxDepth, d = -1;
    uint64_t h = 0;
    uint32_t lo = 0;
    if (e < n) {
        d = dlev[e];
        a = minlev[e] - 1;
        h = khi[e];
        lo = klo[e];
    }
    topo_tile_setup(nlev, ltot, a, d, off, wc);
    if (blockIdx.x == 0 && (int)threadIdx.x <= nlev + 1) lev[threadIdx.x] = off[threadIdx.x];
    int prev = -1, rank_prev = 0;  // this body's box one level up; its rank there
    bool over = false;
    for (int l = 0; l <= nlev; ++l) {
        const bool f = a < l && l <= d;
        const uint64_t b = __ballot(f);
        const int id = topo_rank(l, b, off, wc, tpre, ntiles);
        if (f) {
            if (id < capacity) {
                const int oc = l ? level_digit(h, lo, l) : 0;
                seg_first[id] = e;
                level[id] = l;
                octant[id] = oc;
                bsys[id] = int(lo >> 9);
                atomicAdd(&cnt[kTopoOctCnt + oc], 1);
                if (l == d) atomicAdd(&cnt[kTopoLeafCnt + l], 1);
                // the shallowest box's parent holds e - 1 too: the last box of
                // the level above started before e
                parent[id] = l == 0 ? -1 : (l == a + 1 ? rank_prev - 1 : prev);
                if (l == d) {
                    nchild[id] = 0;
                    child_first[id] = -1;
                    mask[id] = 0;
                }
            } else {
                over = true;
            }
            if (l > a + 1 && prev < capacity) child_first[prev] = id;
            prev = id;
        }
        rank_prev = id;
    }
    if (over) atomicOr(ovf, 1);
    __syncthreads();
    if (threadIdx.x < kTopoMaxLeaf) part[(size_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Box ends, seen from the body after them: for e in [1, n] the boxes of levels
// a(e) < l <= d(e - 1) that hold e - 1 end at e (e = n: every box still open).
// Their body counts, child counts and child octant masks.
__global__ void __launch_bounds__(kTopoTile) k_topo_ends(int n, int nlev, int ntiles, int capacity,
                                                         const int* __restrict__ dlev,
                                                         const int* __restrict__ minlev,
                                                         const int* __restrict__ tpre, const int* __restrict__ ltot,
                                                         const int* __restrict__ seg_first,
                                                         const int* __restrict__ child_first,
                                                         const int* __restrict__ octant, int* seg_count,
                                                         int* nchild, unsigned char* mask, int* __restrict__ part) {
    __shared__ int off[kMaxDepth + 2];
    __shared__ int wc[kMaxDepth + 1][kTopoTile / kWave];
    const int e = blockIdx.x * kTopoTile + threadIdx.x;
    int a = kMaxDepth, d = -1, ae = kMaxDepth, dp = -1;
    if (e < n) {
        d = dlev[e];
        a = minlev[e] - 1;
    }
    if (e >= 1 && e <= n) {
        ae = e < n ? a : -1;
        dp = dlev[e - 1];
    }
    topo_tile_setup(nlev, ltot, a, d, off, wc);
    int rank_prev = 0, big = 0;
    for (int l = 0; l <= nlev + 1; ++l) {
        int r = 0;
        if (l <= nlev) r = topo_rank(l, __ballot(a < l && l <= d), off, wc, tpre, ntiles);
        const int lp = l - 1;
        if (lp >= 0 && ae < lp && lp <= dp) {
            const int P = rank_prev - 1;  // the box of level lp holding e - 1
            if (P < capacity) {
                const int c = e - seg_first[P];
                seg_count[P] = c;
                if (lp == dp) big = max(big, c);  // a leaf
                if (lp < dp && child_first[P] >= 0) {
                    const int cf = child_first[P];
                    nchild[P] = r - cf;
                    unsigned mk = 0;
                    for (int c = cf; c < r && c < capacity; ++c) mk |= 1u << octant[c];
                    mask[P] = (unsigned char)mk;
                }
            }
        }
        rank_prev = r;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) big = max(big, __shfl_xor(big, m, 64));
    __shared__ int wbig[kTopoTile / kWave];
    if ((threadIdx.x & (kWave - 1)) == 0) wbig[threadIdx.x / kWave] = big;
    __syncthreads();
    if (threadIdx.x == 0)
        part[(size_t)kTopoMaxLeaf * ntiles + blockIdx.x] = max(max(wbig[0], wbig[1]), max(wbig[2], wbig[3]));
}

// The counts of the topology passes from their per-block partials (one block
// per count: sums, the largest leaf a maximum) into extra[0, kTopoExtra).
__global__ void __launch_bounds__(kScanThreads) k_topo_sums(int ntiles, const int* __restrict__ part,
                                                            int* __restrict__ extra) {
    __shared__ int lds[kScanThreads / kWave];
    const int k = blockIdx.x;
    const int* row = part + (size_t)k * ntiles;
    int v = 0;
    if (k == kTopoMaxLeaf || k == kTopoDeep) {
        for (int t = threadIdx.x; t < ntiles; t += kScanThreads) v = max(v, row[t]);
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m, 64));
        if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < kScanThreads / kWave; ++w) v = max(v, lds[w]);
            extra[k] = v;
        }
    } else {
        for (int t = threadIdx.x; t < ntiles; t += kScanThreads) v += row[t];
        int total;
        block_exclusive_scan(v, lds, total);
        if (threadIdx.x == 0) extra[k] = total;
    }
}

// Box q of level 0 is the root of system q (bodies [sys_off[q], sys_off[q+1])).
__global__ void k_init_roots(int nsys, const int* sys_off, int* seg_first, int* seg_count, int* level, int* parent,
                             int* octant, int* bsys) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nsys) return;
    seg_first[q] = sys_off[q];
    seg_count[q] = sys_off[q + 1] - sys_off[q];
    level[q] = 0;
    parent[q] = -1;
    octant[q] = 0;
    bsys[q] = q;
}

// ---- reference numbering -------------------------------------------------------
__global__ void k_fill(int* a, int n, int v) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_zero4(int n, int* a, int* b, int* c, int* d) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i < n) a[i] = b[i] = c[i] = d[i] = 0;
}
__global__ void __launch_bounds__(kT) k_count_starts(int nb, const int* seg_first, const int* level,
                                                    int* cnt, int* minlev) {
    const int b = blockIdx.x * kT + threadIdx.x;
    if (b >= nb) return;
    atomicAdd(&cnt[seg_first[b]], 1);
    atomicMin(&minlev[seg_first[b]], level[b]);
}

__global__ void __launch_bounds__(kT) k_preorder(int nb, const int* seg_first, const int* level,
                                                const int* start_scan, const int* minlev,
                                                const int* nchild, int* pre, int* nch_by_pre) {
    const int b = blockIdx.x * kT + threadIdx.x;
    if (b >= nb) return;
    const int s = seg_first[b];
    const int p = start_scan[s] + level[b] - minlev[s];
    pre[b] = p;
    nch_by_pre[p] = nchild[b];
}

// Ids restart at 0 in every system: its root (box q, pre-order rank pre[q])
// precedes all of its boxes, so base_by_pre[pre[q]] counts the boxes of the
// systems before it that are some box's child.
__global__ void __launch_bounds__(kT) k_refid(int nb, const int* pre, const int* base_by_pre, const int* parent,
                                             const int* bsys, const int* child_first, const int* nchild,
                                             int* refid) {
    const int b = blockIdx.x * kT + threadIdx.x;
    if (b >= nb) return;
    if (parent[b] < 0) refid[b] = 0;
    const int k = nchild[b];
    const int first = 1 + base_by_pre[pre[b]] - base_by_pre[pre[bsys[b]]];
    for (int r = 0; r < k; ++r) refid[child_first[b] + r] = first + r;
}

// ---- (level, octant) order and non-leaf lists ------------------------------------
__global__ void __launch_bounds__(kT) k_lo_keys(int nb, const int* level, const int* octant, uint64_t* khi,
                                               uint32_t* klo, int* val) {
    const int b = blockIdx.x * kT + threadIdx.x;
    if (b >= nb) return;
    khi[b] = uint64_t(level[b] * 8 + octant[b]);
    klo[b] = 0;
    val[b] = b;
}
// M2L target order: (octant, level, BFS). A box's separated offsets depend on
// its octant (not its level), so tiles cut inside one octant group share one
// set of at most 189 offsets. It regroups the (level, octant) order: the
// group (l, o) keeps its run, moved from the start the (level, octant) sort
// gave it (digit l*8+o of its scanned histogram, hist[d * nblocks]) to the
// start of (o, l) among the groups ordered by octant first.
constexpr int kLoDigits = 8 * (kMaxDepth + 1);  // (level, octant) keys; < 512 buckets
__global__ void __launch_bounds__(kT) k_m2l_from_lo(int nb, const int* __restrict__ lo_order,
                                                   const int* __restrict__ level, const int* __restrict__ octant,
                                                   const int* __restrict__ hist, int nblocks, int* m2l_order) {
    __shared__ int lo_start[kLoDigits + 1], m2l_start[kLoDigits];
    for (int d = threadIdx.x; d <= kLoDigits; d += kT) lo_start[d] = hist[(size_t)d * nblocks];
    __syncthreads();
    if (threadIdx.x == 0) {
        int pos = 0;
        for (int o = 0; o < 8; ++o)
            for (int l = 0; l <= kMaxDepth; ++l) {
                const int d = l * 8 + o;
                m2l_start[d] = pos;
                pos += lo_start[d + 1] - lo_start[d];
            }
    }
    __syncthreads();
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= nb) return;
    const int b = lo_order[i];
    const int d = level[b] * 8 + octant[b];
    m2l_order[m2l_start[d] + (i - lo_start[d])] = b;
}
// Boxes per octant, added into cnt[0..8) (zeroed before): block sums in LDS,
// then eight integer atomics per block (order-free).
constexpr int kOctCountThreads = 256;
constexpr int kOctCountPer = 16;  // boxes per thread
__global__ void __launch_bounds__(kOctCountThreads) k_octant_count(int nb, const int* octant, int* cnt) {
    __shared__ int c[8];
    if (threadIdx.x < 8) c[threadIdx.x] = 0;
    __syncthreads();
    int my[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int b0 = blockIdx.x * kOctCountThreads * kOctCountPer + (int)threadIdx.x;
#pragma unroll
    for (int u = 0; u < kOctCountPer; ++u) {
        const int b = b0 + u * kOctCountThreads;
        if (b < nb) {
            const int o = octant[b];
#pragma unroll
            for (int q = 0; q < 8; ++q) my[q] += (o == q);
        }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        int v = my[q];
        for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(&c[q], v);
    }
    __syncthreads();
    if (threadIdx.x < 8 && c[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], c[threadIdx.x]);
}
// M2L tile starts over m2l_order (cnt = boxes per octant): each octant group
// split into the fewest near-equal tiles of at most T targets. The host
// derives the tile count from the same counts with m2l_tile_count().
__host__ __device__ inline int m2l_tiles_of(int cnt, int T) { return (cnt + T - 1) / T; }

// Targets per M2L tile for workgroups of up to `tile` targets: the smallest
// size whose tiles still fit one per CU, else `tile`. A workgroup's time is
// set by its busiest SIMD's chain of MFMAs per offset times the offsets, so
// below a CU's worth of full tiles smaller tiles on more CUs finish sooner:
// multiples of 64 targets (one 16-target column block per SIMD), and for the
// 16-wave FP64 workgroups (the 256 plan) also 32 and 16 targets, whose blocks
// the launch splits by rows over two and four waves (shorter chains).
int m2l_tile_size(const int* per_octant, int tile, int cus) {
    auto tiles = [&](int t) {
        int k = 0;
        for (int q = 0; q < 8; ++q)
            if (per_octant[q] > 0) k += m2l_tiles_of(per_octant[q], t);
        return k;
    };
    if (tile == 256)
        for (int t : {16, 32})
            if (tiles(t) <= cus) return t;
    for (int t = 64; t < tile; t += 64)
        if (tiles(t) <= cus) return t;
    return tile;
}

// Compute units of the current device (cached per device).
int device_cus() {
    static std::mutex mu;
    static std::map<int, int> cache;
    int dev = 0;
    AFMM_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(dev);
    if (it != cache.end()) return it->second;
    int n = 0;
    AFMM_CHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    return cache[dev] = n > 0 ? n : 1;
}
__global__ void __launch_bounds__(kT) k_m2l_tile_plan(const int* cnt, int T, int ntiles, int* tstart) {
    __shared__ int c[8], first_tile[9], first_box[9];
    if (threadIdx.x == 0) {
        int k = 0, pos = 0;
        for (int q = 0; q < 8; ++q) {
            c[q] = cnt[q];
            first_tile[q] = k;
            first_box[q] = pos;
            if (c[q] > 0) k += m2l_tiles_of(c[q], T), pos += c[q];
        }
        first_tile[8] = k;
        first_box[8] = pos;
    }
    __syncthreads();
    for (int t = threadIdx.x; t <= ntiles; t += kT) {
        if (t == ntiles) {
            tstart[t] = first_box[8];
            continue;
        }
        int q = 0;
        while (t >= first_tile[q + 1]) ++q;
        const int k = m2l_tiles_of(c[q], T), base = c[q] / k, rem = c[q] % k, i = t - first_tile[q];
        tstart[t] = first_box[q] + i * base + (i < rem ? i : rem);
    }
}
__global__ void __launch_bounds__(kT) k_nonleaf_ids(int nb, const int* nchild, const int* leaf_scan, int* ids) {
    const int b = blockIdx.x * kT + threadIdx.x;
    if (b < nb && nchild[b] > 0) ids[b - leaf_scan[b]] = b;
}
// leaf_scan at each level boundary (levels given by host offsets)
// out[i] = a[idx[i]] (-1 past total) for i < n, and out[n + 8 + k] = extra[k]
// for k < nextra (the 8 entries between are the octant counts).
__global__ void k_gather_at(const int* a, const int* idx, int n, int total, int* out, const int* extra, int nextra) {
    const int i = threadIdx.x;
    if (i < n) out[i] = idx[i] < total ? a[idx[i]] : -1;
    if (i < 8) out[n + i] = 0;  // the octant counts, summed next
    if (i < nextra) out[n + 8 + i] = extra[i];
}

// ---- leaves and the in-leaf input order ------------------------------------------
// Leaf flags, and the largest leaf's body count into *maxleaf (zeroed before;
// null: not wanted).
__global__ void __launch_bounds__(kT) k_leaf_flags(int nb, const int* nchild, const int* seg_count, int* flag,
                                                  int* maxleaf) {
    const int b = blockIdx.x * kT + threadIdx.x;
    int c = 0;
    if (b < nb) {
        const bool leaf = nchild[b] == 0;
        flag[b] = leaf;
        if (leaf) c = seg_count[b];
    }
    if (!maxleaf) return;
    for (int m = 32; m >= 1; m >>= 1) c = max(c, __shfl_xor(c, m, 64));
    if ((threadIdx.x & 63) == 0 && c > 0) atomicMax(maxleaf, c);
}
__global__ void __launch_bounds__(kT) k_leaf_ids(int nb, const int* nchild, const int* scan, int* ids) {
    const int b = blockIdx.x * kT + threadIdx.x;
    if (b < nb && nchild[b] == 0) ids[scan[b]] = b;
}
// key2[input index] = first sorted slot of its leaf; leaf_of[sorted slot] = leaf box.
// One wave per leaf (lanes over its bodies).
__global__ void __launch_bounds__(kT) k_leaf_keys(int nleaves, const int* leaf_ids, const int* seg_first,
                                                 const int* seg_count, const int* perm1, uint64_t* key2,
                                                 uint32_t* klo2, int* val2, int* leaf_of) {
    const int li = blockIdx.x * (kT / 64) + (int)(threadIdx.x >> 6);
    if (li >= nleaves) return;
    const int b = leaf_ids[li];
    const int f = seg_first[b], c = seg_count[b];
    for (int k = f + (int)(threadIdx.x & 63); k < f + c; k += 64) {
        const int j = perm1[k];
        key2[j] = (uint64_t)f;
        klo2[j] = 0;
        val2[j] = j;
        leaf_of[k] = b;
    }
}
constexpr int kLeafRankMax = 2048;  // per-leaf ranking is O(c^2 / 64) per lane

// Input order inside every leaf: the final order of leaf b's slots
// [f, f + c) is its input indices perm1[f .. f + c) ascending (what a stable
// sort of all bodies by leaf start would give), here ranked inside the leaf by
// one wave (indices are distinct). Also leaf_of[slot] = b.
// The ranked body is gathered into its slot at once (positions, mass, input
// index), in place of a separate pass over the permutation.
__global__ void __launch_bounds__(kT) k_leaf_order(int nleaves, const int* leaf_ids, const int* seg_first,
                                                  const int* seg_count, const int* __restrict__ perm1,
                                                  const double* __restrict__ pos, const double* __restrict__ mass,
                                                  double* __restrict__ x, double* __restrict__ y,
                                                  double* __restrict__ z, double* __restrict__ m,
                                                  int* __restrict__ orig, int* __restrict__ leaf_of) {
    const int li = blockIdx.x * (kT / 64) + (int)(threadIdx.x >> 6);
    if (li >= nleaves) return;
    const int b = leaf_ids[li];
    const int f = seg_first[b], c = seg_count[b];
    for (int k = (int)(threadIdx.x & 63); k < c; k += 64) {
        const int v = perm1[f + k];
        const double px = pos[3 * (size_t)v], py = pos[3 * (size_t)v + 1], pz = pos[3 * (size_t)v + 2];
        const double pm = mass[v];
        int r = 0;
        for (int j = 0; j < c; ++j) r += perm1[f + j] < v;
        x[f + r] = px;
        y[f + r] = py;
        z[f + r] = pz;
        m[f + r] = pm;
        orig[f + r] = v;
        leaf_of[f + k] = b;
    }
}

__global__ void __launch_bounds__(kT) k_gather(int n, const int* perm, const double* pos, const double* mass,
                                              double* x, double* y, double* z, double* m, int* orig) {
    const int k = blockIdx.x * kT + threadIdx.x;
    if (k >= n) return;
    const int j = perm[k];
    x[k] = pos[3 * (size_t)j];
    y[k] = pos[3 * (size_t)j + 1];
    z[k] = pos[3 * (size_t)j + 2];
    m[k] = mass[j];
    orig[k] = j;
}

// ---- box geometry -----------------------------------------------------------------
// Replays the reference's centre arithmetic along the box's path (digits of
// any of its bodies' keys), and derives the exact centre and integer coords.
__global__ void __launch_bounds__(kT) k_geometry(int nb, const RootInfo* roots, const int* bsys,
                                                const int* seg_first, const int* level, const uint64_t* khi,
                                                const uint32_t* klo, double* cfp, double* cdd, double* half,
                                                double* invh, int* ic) {
    const int b = blockIdx.x * kT + threadIdx.x;
    if (b >= nb) return;
    const RootInfo r = roots[bsys[b]];
    const int L = level[b];
    const uint64_t kh = khi[seg_first[b]];
    const uint32_t kl = klo[seg_first[b]];
    double c[3] = {r.c[0], r.c[1], r.c[2]}, h = r.half;
    int ii[3] = {0, 0, 0};
    for (int l = 1; l <= L; ++l) {
        const int o = level_digit(kh, kl, l);
        h = 0.5 * h;
        for (int a = 0; a < 3; ++a) {
            const int bit = (o >> a) & 1;
            c[a] = c[a] + (bit ? h : -h);
            ii[a] = 2 * ii[a] + bit;
        }
    }
    for (int a = 0; a < 3; ++a) {
        cfp[3 * (size_t)b + a] = c[a];
        const Centre e = exact_centre(r.c[a], r.half, L, ii[a]);
        cdd[6 * (size_t)b + 2 * a] = e.hi;
        cdd[6 * (size_t)b + 2 * a + 1] = e.lo;
        ic[3 * (size_t)b + a] = ii[a];
    }
    half[b] = h;
    invh[b] = ldexp(1.0 / r.half, L);
}

// ---- lists ------------------------------------------------------------------------
__device__ __forceinline__ bool adjacent(const double* cfp, const double* half, double tol, int a, int b) {
    const double reach = half[a] + half[b] + tol;
    return fabs(cfp[3 * (size_t)a] - cfp[3 * (size_t)b]) <= reach &&
           fabs(cfp[3 * (size_t)a + 1] - cfp[3 * (size_t)b + 1]) <= reach &&
           fabs(cfp[3 * (size_t)a + 2] - cfp[3 * (size_t)b + 2]) <= reach;
}

// Colleagues and V of every box of one level (AdaptiveOctree.hpp:143-161):
// children of the parent's colleagues, colleague order then octant order.
// One wave per box: lane k takes candidate k = 8*slot + octant (<= 216), and
// ballots keep the emitted order (= ascending id = Morton order in a level).
// The same pass carries, level by level, the two upward-looking sets the U,
// W and X lists are built from (AdaptiveOctree.hpp:163-189):
//  * A(b) = the leaves of b's level or coarser that are adjacent to b: a leaf
//    colleague of b (b itself when b is a leaf), or an entry of A(parent)
//    adjacent to b. At most one such leaf covers each of the 27 cells around
//    b, so |A(b)| <= 27. For a leaf b, A(b) is the coarser and same-level part
//    of U(b).
//  * X(b) = the entries of A(parent) not adjacent to b. An entry c of
//    A(parent) is a leaf colleague of an ancestor A' of b, and every box
//    between A' and b's parent is adjacent to c: descendW from c enters them
//    (adjacent non-leaves) and stops at b, which it puts in W(c) because b is
//    separated from c. So X(b) is exactly {c : b in W(c)}, the reference's X
//    (:172-173), ranked here by reference id, the order it is filled in.
// The finer part of U and all of W are the transposes: b in U(c) for every
// leaf b with c in A(b) coarser than b, and b in W(c) for every c in X(b).
__global__ void __launch_bounds__(64) k_coll_v(int off, const RootInfo* __restrict__ roots,
                                              const int* __restrict__ bsys, const int* __restrict__ parent,
                                              const int* __restrict__ child_first, const int* __restrict__ nchild,
                                              const unsigned char* __restrict__ mask, const double* __restrict__ cfp,
                                              const double* __restrict__ half, const int* __restrict__ ic,
                                              const int* __restrict__ refid,
                                              int* __restrict__ coll, int* __restrict__ ncoll, int* __restrict__ vl,
                                              int* __restrict__ nv, int* __restrict__ aset,
                                              int* __restrict__ acnt, int* __restrict__ xset, int* __restrict__ xcnt,
                                              int* __restrict__ overflow) {
    __shared__ int xbox[kCollSlots], xkey[kCollSlots];
    const int b = off + blockIdx.x;
    const int lane = threadIdx.x;
    const int par = parent[b];
    if (par < 0) {  // a root: colleague[root] = {root} (AdaptiveOctree.hpp:143)
        if (lane == 0) {
            coll[kCollSlots * (size_t)b] = b;
            ncoll[b] = 1;
            nv[b] = 0;
            const bool leaf = nchild[b] == 0;
            if (leaf) aset[kCollSlots * (size_t)b] = b;
            acnt[b] = leaf ? 1 : 0;
            xcnt[b] = 0;
        }
        return;
    }
    const double tol = roots[bsys[b]].tol;
    const int npc = ncoll[par];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // All candidates' loads first (lane k of round r takes candidate
    // 64 r + k = 8 slot + octant, at most 8 * 27 = 216), so that the four
    // rounds' dependent load chains overlap; the ballots then place them in
    // candidate order. Also entry `lane` of A(parent) -> A(b) or X(b).
    constexpr int R = (8 * kCollSlots + 63) / 64;
    int cand[R];
    bool cadj[R], cleaf[R], cfar[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int k = 64 * r + lane;
        const int slot = k >> 3, o = k & 7;
        cand[r] = -1;
        if (slot < npc) {
            const int pc = coll[kCollSlots * (size_t)par + slot];
            const unsigned mk = mask[pc];
            if ((mk >> o) & 1u) cand[r] = child_first[pc] + __popc(mk & ((1u << o) - 1u));
        }
    }
    int pa = -1, pa_ref = 0;
    bool pa_adj = false;
    if (lane < acnt[par]) pa = aset[kCollSlots * (size_t)par + lane];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int c = cand[r];
        cadj[r] = cleaf[r] = cfar[r] = false;
        if (c >= 0) {
            cadj[r] = adjacent(cfp, half, tol, b, c);
            cleaf[r] = nchild[c] == 0;
            const int dx = ic[3 * (size_t)b] - ic[3 * (size_t)c], dy = ic[3 * (size_t)b + 1] - ic[3 * (size_t)c + 1],
                      dz = ic[3 * (size_t)b + 2] - ic[3 * (size_t)c + 2];
            cfar[r] = dx < -3 || dx > 3 || dy < -3 || dy > 3 || dz < -3 || dz > 3;
        }
    }
    if (pa >= 0) {
        pa_adj = adjacent(cfp, half, tol, pa, b);
        pa_ref = refid[pa];
    }
    const uint64_t ba = __ballot(pa >= 0 && pa_adj), bx = __ballot(pa >= 0 && !pa_adj);
    if (pa >= 0 && pa_adj) aset[kCollSlots * (size_t)b + __popcll(ba & lt)] = pa;  // < acnt[par] <= 27
    if (pa >= 0 && !pa_adj) {
        xbox[__popcll(bx & lt)] = pa;
        xkey[__popcll(bx & lt)] = pa_ref;
    }
    int ka = __popcll(ba);
    const int kx = __popcll(bx);
    int kc = 0, kv = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int c = cand[r];
        const bool adj = cadj[r], leafc = c >= 0 && adj && cleaf[r];
        const uint64_t bc = __ballot(c >= 0 && adj), bv = __ballot(c >= 0 && !adj), bl = __ballot(leafc);
        if (c >= 0 && adj) {
            const int pos = kc + __popcll(bc & lt);
            if (pos < kCollSlots) coll[kCollSlots * (size_t)b + pos] = c;
            else atomicOr(overflow, 1);
        } else if (c >= 0) {
            const int pos = kv + __popcll(bv & lt);
            if (pos < kVSlots) vl[kVSlots * (size_t)b + pos] = c;
            else atomicOr(overflow, 1);
            if (cfar[r]) atomicOr(overflow, 2);
        }
        if (leafc) {
            const int pos = ka + __popcll(bl & lt);
            if (pos < kCollSlots) aset[kCollSlots * (size_t)b + pos] = c;
            else atomicOr(overflow, 1);
        }
        kc += __popcll(bc);
        kv += __popcll(bv);
        ka += __popcll(bl);
    }
    // X(b) ranked by reference id (ids are distinct)
    __syncthreads();
    if (lane < kx) {
        const int key = xkey[lane];
        int r = 0;
        for (int j = 0; j < kx; ++j) r += xkey[j] < key;
        xset[kCollSlots * (size_t)b + r] = xbox[lane];
    }
    if (lane == 0) {
        ncoll[b] = kc < kCollSlots ? kc : kCollSlots;
        nv[b] = kv < kVSlots ? kv : kVSlots;
        acnt[b] = ka < kCollSlots ? ka : kCollSlots;
        xcnt[b] = kx;
    }
}

__device__ __forceinline__ int wave_total(int v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// ---- U, W and X from the A and X sets -------------------------------------------
// List sizes, one thread per (box, slot) of the A and X sets: |U(b)| =
// |A(b)| plus one for every finer leaf that has b in its A set; |W(c)| = the
// boxes whose X holds c. Integer atomics: the counts do not depend on the order.
__global__ void __launch_bounds__(kT) k_list_count(int nb, const int* __restrict__ nchild,
                                                  const int* __restrict__ level, const int* __restrict__ leaf_scan,
                                                  const int* __restrict__ aset, const int* __restrict__ acnt,
                                                  const int* __restrict__ xset, const int* __restrict__ xcnt,
                                                  int* __restrict__ ucnt, int* __restrict__ wcnt) {
    const long long g = (long long)blockIdx.x * kT + threadIdx.x;
    if (g >= (long long)nb * kCollSlots) return;
    const int b = (int)(g / kCollSlots), k = (int)(g - (long long)b * kCollSlots);
    if (nchild[b] == 0) {
        const int na = acnt[b];
        if (k == 0) atomicAdd(&ucnt[leaf_scan[b]], na);
        if (k < na) {
            const int c = aset[g];
            if (level[c] < level[b]) atomicAdd(&ucnt[leaf_scan[c]], 1);
        }
    }
    if (k < xcnt[b]) atomicAdd(&wcnt[leaf_scan[xset[g]]], 1);
}

// Fills the lists, one thread per (box, slot): U(b) starts with A(b); the
// transposed entries (finer U, W) take slots from per-list cursors (zeroed
// before), in an order k_list_sort then fixes. X(b) keeps its ranked order.
__global__ void __launch_bounds__(kT) k_list_fill(int nb, const int* __restrict__ nchild,
                                                 const int* __restrict__ level, const int* __restrict__ leaf_scan,
                                                 const int* __restrict__ aset, const int* __restrict__ acnt,
                                                 const int* __restrict__ xset, const int* __restrict__ xcnt,
                                                 const int* __restrict__ uoff, const int* __restrict__ woff,
                                                 const int* __restrict__ xoff, int* __restrict__ ucur,
                                                 int* __restrict__ wcur, int* __restrict__ uidx,
                                                 int* __restrict__ widx, int* __restrict__ xidx) {
    const long long g = (long long)blockIdx.x * kT + threadIdx.x;
    if (g >= (long long)nb * kCollSlots) return;
    const int b = (int)(g / kCollSlots), k = (int)(g - (long long)b * kCollSlots);
    if (nchild[b] == 0 && k < acnt[b]) {
        const int c = aset[g];
        uidx[uoff[leaf_scan[b]] + k] = c;
        if (level[c] < level[b]) {
            const int lc = leaf_scan[c];
            uidx[uoff[lc] + acnt[c] + atomicAdd(&ucur[lc], 1)] = b;
        }
    }
    if (k < xcnt[b]) {
        const int c = xset[g];
        xidx[xoff[b] + k] = c;
        const int lc = leaf_scan[c];
        widx[woff[lc] + atomicAdd(&wcur[lc], 1)] = b;
    }
}

// Puts U and W of each leaf in the reference's order, one wave per leaf: U
// in DFS order from the root, W in colleague order then DFS order
// (AdaptiveOctree.hpp:163-189). The entries of each list are disjoint boxes,
// whose DFS order is the order of their first bodies, and colleagues come in
// Morton order, which is again body order, so both are ranks by first body.
// Also the P2P source count of the leaf (bodies in its U leaves).
constexpr int kListSortCap = 512;
__device__ void sort_by_first_body(int* __restrict__ d, int n, const int* __restrict__ seg_first, int* box,
                                   int* key, int lane) {
    if (n <= 1) return;
    if (n <= kListSortCap) {
        for (int i = lane; i < n; i += 64) {
            box[i] = d[i];
            key[i] = seg_first[box[i]];
        }
        __syncthreads();
        for (int i = lane; i < n; i += 64) {
            const int k = key[i];
            int r = 0;
            for (int j = 0; j < n; ++j) r += key[j] < k;
            d[r] = box[i];
        }
        __syncthreads();
    } else if (lane == 0) {
        for (int i = 1; i < n; ++i) {
            const int v = d[i], k = seg_first[v];
            int j = i - 1;
            while (j >= 0 && seg_first[d[j]] > k) {
                d[j + 1] = d[j];
                --j;
            }
            d[j + 1] = v;
        }
    }
}
__global__ void __launch_bounds__(64) k_list_sort(const int* __restrict__ uoff, const int* __restrict__ woff,
                                                 const int* __restrict__ seg_first,
                                                 const int* __restrict__ seg_count, int* __restrict__ uidx,
                                                 int* __restrict__ widx, int* __restrict__ usrc) {
    __shared__ int box[kListSortCap], key[kListSortCap];
    const int li = blockIdx.x, lane = threadIdx.x;
    const int u0 = uoff[li], nu = uoff[li + 1] - u0;
    int ns = 0;
    for (int i = lane; i < nu; i += 64) ns += seg_count[uidx[u0 + i]];
    ns = wave_total(ns);
    if (lane == 0) usrc[li] = ns;
    sort_by_first_body(uidx + u0, nu, seg_first, box, key, lane);
    const int w0 = woff[li];
    sort_by_first_body(widx + w0, woff[li + 1] - w0, seg_first, box, key, lane);
}

// The three list sizes and the slot-overflow flag, for one readback.
__global__ void k_list_sizes(const int* nu, const int* nw, const int* nx, const int* overflow, int* out,
                             unsigned long long* totals) {
    if (threadIdx.x < 4) totals[threadIdx.x] = 0ull;  // zeroed for k_list_totals
    if (threadIdx.x == 0) {
        out[0] = *nu;
        out[1] = *nw;
        out[2] = *nx;
        out[3] = *overflow;
    }
}

// Totals for the statistics: colleague and V entries, the largest leaf and
// sum_leaf n_leaf * usrc_leaf: block sums in LDS, then four integer atomics
// per block (order-free).
constexpr int kTotThreads = 1024;
__global__ void __launch_bounds__(kTotThreads) k_list_totals(int nb, int nleaves, const int* __restrict__ ncoll,
                                                            const int* __restrict__ nv,
                                                            const int* __restrict__ leaf_ids,
                                                            const int* __restrict__ seg_count,
                                                            const int* __restrict__ usrc,
                                                            unsigned long long* __restrict__ tot) {
    __shared__ unsigned long long red[4][kTotThreads / 64];
    unsigned long long c = 0, v = 0, occ = 0, pairs = 0;
    const int i = blockIdx.x * kTotThreads + threadIdx.x;
    if (i < nb) {
        c = (unsigned long long)ncoll[i];
        v = (unsigned long long)nv[i];
    }
    if (i < nleaves) {
        occ = (unsigned long long)seg_count[leaf_ids[i]];
        pairs = occ * (unsigned long long)usrc[i];
    }
    for (int m = 32; m >= 1; m >>= 1) {
        c += __shfl_xor(c, m, 64);
        v += __shfl_xor(v, m, 64);
        pairs += __shfl_xor(pairs, m, 64);
        const unsigned long long o = __shfl_xor(occ, m, 64);
        occ = o > occ ? o : occ;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = c;
        red[1][w] = v;
        red[2][w] = occ;
        red[3][w] = pairs;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        unsigned long long t = 0;
        for (int k = 0; k < kTotThreads / 64; ++k) {
            const unsigned long long x = red[threadIdx.x][k];
            t = threadIdx.x == 2 ? (x > t ? x : t) : t + x;
        }
        if (threadIdx.x == 2) atomicMax(&tot[2], t);
        else atomicAdd(&tot[threadIdx.x], t);
    }
}

int read_int(const int* d, hipStream_t s) {
    int h = 0;
    AFMM_CHECK(hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, s));
    AFMM_CHECK(hipStreamSynchronize(s));
    return h;
}

}  // namespace

void build_forest(Pool& pool, hipStream_t s, const double* d_pos, const double* d_mass, int nsys,
                  const int* h_sys_off, const double* h_eps_floor, int cap, TreeView& tv, TreeMeta& meta,
                  std::vector<RootInfo>& roots, const TreeHook& on_core, const TreeHook& on_vlists,
                  hipStream_t ls, hipEvent_t ls_ready, int* levels_hint, int* depth_hint) {
    const int n = h_sys_off[nsys];
    // ---- per-system root cubes: chunks of at most kBBoxChunk bodies of one system
    std::vector<int> hb;  // [cbeg | cend | cfirst(nsys+1) | sys_off(nsys+1)] as one upload
    std::vector<int> cb, ce, cf(nsys + 1, 0);
    for (int q = 0; q < nsys; ++q) {
        cf[q] = (int)cb.size();
        for (int i = h_sys_off[q]; i < h_sys_off[q + 1]; i += kBBoxChunk) {
            cb.push_back(i);
            ce.push_back(std::min(i + kBBoxChunk, h_sys_off[q + 1]));
        }
    }
    cf[nsys] = (int)cb.size();
    const int nch = (int)cb.size();
    hb.insert(hb.end(), cb.begin(), cb.end());
    hb.insert(hb.end(), ce.begin(), ce.end());
    hb.insert(hb.end(), cf.begin(), cf.end());
    hb.insert(hb.end(), h_sys_off, h_sys_off + nsys + 1);
    int* d_hb = pool.get<int>("sys_ints", hb.size());
    double* d_eps_floor = pool.get<double>("sys_eps_floor", nsys);
    AFMM_CHECK(hipMemcpyAsync(d_hb, hb.data(), sizeof(int) * hb.size(), hipMemcpyHostToDevice, s));
    AFMM_CHECK(hipMemcpyAsync(d_eps_floor, h_eps_floor, sizeof(double) * nsys, hipMemcpyHostToDevice, s));
    const int* d_cbeg = d_hb;
    const int* d_cend = d_hb + nch;
    const int* d_cfirst = d_hb + 2 * nch;
    const int* d_sys_off = d_hb + 2 * nch + nsys + 1;
    double* pv = pool.get<double>("bbox_v", 6 * (size_t)nch);
    int* pi = pool.get<int>("bbox_i", 6 * (size_t)nch);
    int* pbad = pool.get<int>("bbox_bad", nch);
    RootInfo* d_roots = pool.get<RootInfo>("roots", nsys);
    hipLaunchKernelGGL(k_bbox_partial, dim3(nch), dim3(kT), 0, s, d_pos, d_mass, d_cbeg, d_cend, pv, pi, pbad);
    hipLaunchKernelGGL(k_bbox_final, dim3(nsys), dim3(kT), 0, s, pv, pi, pbad, nch, d_cfirst, d_eps_floor, d_roots);
    // (into pinned memory: a copy to pageable memory would wait for the GPU)
    RootInfo* h_roots = pool.host<RootInfo>("h_roots", nsys);
    AFMM_CHECK(hipMemcpyAsync(h_roots, d_roots, sizeof(RootInfo) * nsys, hipMemcpyDeviceToHost, s));
    // The key depth L (levels that may split, the deepest over the systems)
    // sets the number of sort passes and of topology levels. With the
    // previous evaluation's depth as a guess, the root cubes come back with
    // the level offsets (one round trip fewer); a deeper actual depth sorts
    // and builds again, a shallower one gives the same tree (the extra key
    // digits are zero and every system stops at its own depth).
    const int hint = levels_hint ? *levels_hint : -1;
    bool guess = hint > 0 && cap <= kTopoWinMax && (size_t)(n + 1) * (hint + 1) < (size_t)INT_MAX / 2;
    auto roots_back = [&]() {  // after a synchronisation: the depth, or -1 for a NaN/Inf input
        roots.assign(h_roots, h_roots + nsys);
        int d = 0;
        for (const RootInfo& r : roots) {
            if (r.nonfinite) return -1;
            d = std::max(d, r.key_levels);
        }
        if (levels_hint) *levels_hint = d;
        return d;
    };
    int L = hint;
    if (!guess) {
        AFMM_CHECK(hipStreamSynchronize(s));  // also retires the uploads of hb
        L = roots_back();
        if (L < 0) return;
    }
    tv.n = n;
    tv.nsys = nsys;
    tv.roots = d_roots;

    // ---- pass 1: sort bodies by octant path
    const int nsb = sort_blocks(n);
    SortBufs sb;
    for (int k = 0; k < 2; ++k) {
        sb.khi[k] = pool.get<uint64_t>(k ? "khi1" : "khi0", n);
        sb.klo[k] = pool.get<uint32_t>(k ? "klo1" : "klo0", n);
        sb.val[k] = pool.get<int>(k ? "val1" : "val0", n);
    }
    sb.hist = pool.get<int>("sort_hist", 512 * (size_t)nsb);
    int scan_cap = std::max(n, 512 * nsb) + 1;
    int* scan_tmp = pool.get<int>("scan_tmp", 2 * scan_tmp_ints(scan_cap) + 16);
    sb.scan_tmp = scan_tmp;
    // the pass-1 keys and permutation stay where the sort leaves them (nothing
    // below sorts with these buffers again)
    const uint64_t* khi = nullptr;
    const uint32_t* klo = nullptr;
    const int* perm1 = nullptr;
    // Levels the sort orders (three per 9-bit digit): every key level, or, with
    // the depth of the context's previous tree as a guess, the levels down to
    // it. The topology only looks that deep, the in-leaf order is the input
    // order whatever the deeper digits, and a body whose leaf lies deeper
    // sends the build round again with all levels.
    const int dguess = depth_hint && *depth_hint > 0 ? *depth_hint : kMaxDepth;
    int D = 0;
    auto sorted_levels = [&]() { D = std::min(L, 3 * ((std::min(dguess, L) + 2) / 3)); };
    sorted_levels();
    auto sort_bodies = [&]() {
        hipLaunchKernelGGL(k_keys, dim3(blocks(n)), dim3(kT), 0, s, d_pos, n, nsys, d_sys_off, d_roots, sb.khi[0],
                           sb.klo[0], sb.val[0]);
        const int ndig = (D + 2) / 3;
        // system digits (klo bits 9..26) after the path digits: sorted by (system, path)
        const int sdig = nsys <= 1 ? 0 : (nsys <= 512 ? 1 : 2);
        // one system with at most 21 levels: every klo is zero and is not moved
        const bool lo = nsys > 1 || L > 21;
        int r1 = radix_sort(sb, n, 8 - ndig, ndig, s, 0, lo);
        if (sdig > 0) r1 = radix_sort(sb, n, 8, sdig, s, r1);
        khi = sb.khi[r1];
        klo = lo ? sb.klo[r1] : sb.klo[0];
        perm1 = sb.val[r1];
    };
    sort_bodies();

    // ---- topology, level by level
    size_t cap_boxes = std::max<size_t>(1024, (size_t)n / 4 + 64);
    auto box_arrays = [&](size_t c, bool keep) {
        tv.seg_first = pool.get<int>("seg_first", c, s, keep);
        tv.seg_count = pool.get<int>("seg_count", c, s, keep);
        tv.level = pool.get<int>("level", c, s, keep);
        tv.parent = pool.get<int>("parent", c, s, keep);
        tv.octant = pool.get<int>("octant", c, s, keep);
        tv.child_first = pool.get<int>("child_first", c, s, keep);
        tv.nchild = pool.get<int>("nchild", c, s, keep);
        tv.child_mask = pool.get<unsigned char>("child_mask", c, s, keep);
        tv.bsys = pool.get<int>("bsys", c, s, keep);
    };
    int* d_total = pool.get<int>("total", 4);
    // boxes started per body and the shallowest level of them (the
    // reference-id pass below): written by the count pass, else counted there
    int* cnt_start = pool.get<int>("cnt_start", n + 1);
    int* minlev = pool.get<int>("minlev", n + 1);
    bool starts_ready = false;
    int l = 0;
    bool done = false;
    std::vector<int> topo_extra;       // leaves per level, boxes per octant, largest leaf
    const int* d_oct_cnt = nullptr;    // boxes per octant on the device (the M2L tile plan)
    // Levels 0..L-1 may split. All levels come from the sorted keys at once
    // (four launches, one read-back of the level offsets) when the leaf
    // capacity fits the window of the count pass; the box arrays are sized
    // for the worst case where that bound is modest. Otherwise, or when the
    // boxes overflow the arrays, the loop below builds one level at a time
    // and reads each level's size back.
    for (;;) {  // again when a guessed key depth or sorted depth was too shallow
    size_t bound = (size_t)nsys;
    for (int lv = 0; lv < L; ++lv) {
        const double full = (double)nsys * std::pow(8.0, lv);
        bound += 8 * (size_t)std::min<double>(full, (double)n / (cap + 1) + 1.0);
    }
    const bool topo_levels = L > 0 && cap <= kTopoWinMax && (size_t)(n + 1) * (L + 1) < (size_t)INT_MAX / 2;
    cap_boxes = std::max(cap_boxes, (size_t)nsys + 1024);
    if (topo_levels && bound <= (size_t)8 * n + (1u << 20) && bound < (size_t)INT_MAX / 2)
        cap_boxes = std::max(cap_boxes, bound);
    box_arrays(cap_boxes, false);
    if (topo_levels) {
        const int ntiles = (n + 1 + kTopoTile - 1) / kTopoTile;  // bodies 0..n (n: the ends of the last boxes)
        // levels 0..D (the sorted ones): offsets [0, D + 1], the overflow flag, then the counts
        int* lev = pool.get<int>("level_dev", D + 3 + kTopoExtra);
        int* ovf = lev + D + 2;
        int* extra = lev + D + 3;
        int* dl = pool.get<int>("topo_leaf_level", n + 1);
        int* tcnt = pool.get<int>("topo_tiles", (size_t)(D + 1) * ntiles);
        int* ltot = pool.get<int>("topo_totals", D + 1);
        int* part = pool.get<int>("topo_partials", (size_t)kTopoExtra * ntiles);
        hipLaunchKernelGGL(k_topo_count, dim3(ntiles), dim3(kTopoTile), 0, s, n, cap, D, ntiles, d_roots, khi, klo,
                           dl, minlev, cnt_start, tcnt, ovf, part);
        hipLaunchKernelGGL(k_topo_scan, dim3(D + 1), dim3(kScanThreads), 0, s, ntiles, tcnt, ltot);
        hipLaunchKernelGGL(k_topo_emit, dim3(ntiles), dim3(kTopoTile), 0, s, n, D, ntiles, (int)cap_boxes, khi, klo,
                           dl, minlev, tcnt, ltot, lev, ovf, tv.seg_first, tv.level, tv.parent, tv.octant,
                           tv.child_first, tv.nchild, tv.child_mask, tv.bsys, part);
        hipLaunchKernelGGL(k_topo_ends, dim3(ntiles), dim3(kTopoTile), 0, s, n, D, ntiles, (int)cap_boxes, dl, minlev,
                           tcnt, ltot, tv.seg_first, tv.child_first, tv.octant, tv.seg_count, tv.nchild,
                           tv.child_mask, part);
        hipLaunchKernelGGL(k_topo_sums, dim3(kTopoExtra), dim3(kScanThreads), 0, s, ntiles, part, extra);
        int* h_lev = pool.host<int>("h_level_dev", D + 3 + kTopoExtra);
        AFMM_CHECK(hipMemcpyAsync(h_lev, lev, sizeof(int) * (D + 3 + kTopoExtra), hipMemcpyDeviceToHost, s));
        AFMM_CHECK(hipStreamSynchronize(s));  // (with a guessed depth also the root cubes and hb)
        const std::vector<int> hl(h_lev, h_lev + D + 3 + kTopoExtra);
        if (guess) {
            const int actual = roots_back();
            if (actual < 0) return;
            guess = false;
            if (actual > L) {
                L = actual;
                sorted_levels();
                sort_bodies();
                continue;
            }
        }
        if (hl[D + 3 + kTopoDeep] && D < L) {  // a leaf below the sorted levels: all levels
            D = L;
            sort_bodies();
            continue;
        }
        if (hl[D + 2] == 0) {
            topo_extra.assign(hl.begin() + D + 3, hl.end());
            d_oct_cnt = extra + kTopoOctCnt;
            meta.level_off.assign(hl.begin(), hl.begin() + D + 2);
            while (meta.level_off.size() > 2 &&
                   meta.level_off[meta.level_off.size() - 1] == meta.level_off[meta.level_off.size() - 2])
                meta.level_off.pop_back();
            l = (int)meta.level_off.size() - 2;
            done = true;
            starts_ready = true;
        }
    }
    break;
    }
    tv.tol = roots[0].tol;
    if (!done) {
    hipLaunchKernelGGL(k_init_roots, dim3((nsys + 255) / 256), dim3(256), 0, s, nsys, d_sys_off, tv.seg_first,
                       tv.seg_count, tv.level, tv.parent, tv.octant, tv.bsys);
    meta.level_off.assign(1, 0);
    meta.level_off.push_back(nsys);
    l = 0;
    for (;; ++l) {
        const int off = meta.level_off[l], cnt = meta.level_off[l + 1] - off;
        int* bnd = pool.get<int>("split_bnd", 9 * (size_t)cnt);
        int* nch = pool.get<int>("split_nch", cnt);
        int* cbase = pool.get<int>("split_base", cnt);
        hipLaunchKernelGGL(k_split, dim3(blocks(8LL * cnt)), dim3(kT), 0, s, off, cnt, l, cap, d_eps_floor, d_roots,
                           tv.bsys, tv.seg_first, tv.seg_count, khi, klo, bnd, nch);
        scan_exclusive(nch, cbase, cnt, d_total, scan_tmp, s);
        const int next = read_int(d_total, s);
        const int next_off = off + cnt;
        if ((size_t)(next_off + next) > cap_boxes) {
            cap_boxes = std::max(cap_boxes * 2, (size_t)(next_off + next) + 1024);
            box_arrays(cap_boxes, true);
        }
        hipLaunchKernelGGL(k_emit, dim3(blocks(cnt)), dim3(kT), 0, s, off, cnt, l, next_off, bnd, nch, cbase,
                           tv.seg_first, tv.seg_count, tv.level, tv.parent, tv.octant, tv.child_first,
                           tv.nchild, tv.child_mask, tv.bsys);
        if (next == 0) break;
        meta.level_off.push_back(next_off + next);
    }
    }
    tv.max_level = l;
    if (depth_hint) *depth_hint = l;
    const int nb = meta.level_off.back();
    tv.nb = nb;
    // scans over boxes from here on: chains of single-child boxes (near
    // duplicates, small capacities) can hold more boxes than bodies
    if (nb + 1 > scan_cap) {
        scan_cap = nb + 1;
        scan_tmp = pool.get<int>("scan_tmp", 2 * scan_tmp_ints(scan_cap) + 16);
        sb.scan_tmp = scan_tmp;
    }

    // ---- reference box ids
    int* start_scan = pool.get<int>("start_scan", n + 1);
    int* pre = pool.get<int>("pre", nb);
    int* nch_pre = pool.get<int>("nch_pre", nb);
    int* base_pre = pool.get<int>("base_pre", nb);
    tv.refid = pool.get<int>("refid", nb);
    if (!starts_ready) {
        hipLaunchKernelGGL(k_fill, dim3(blocks(n)), dim3(kT), 0, s, cnt_start, n, 0);
        hipLaunchKernelGGL(k_fill, dim3(blocks(n)), dim3(kT), 0, s, minlev, n, INT_MAX);
        hipLaunchKernelGGL(k_count_starts, dim3(blocks(nb)), dim3(kT), 0, s, nb, tv.seg_first, tv.level, cnt_start,
                           minlev);
    }
    scan_exclusive(cnt_start, start_scan, n, nullptr, scan_tmp, s);
    hipLaunchKernelGGL(k_preorder, dim3(blocks(nb)), dim3(kT), 0, s, nb, tv.seg_first, tv.level, start_scan,
                       minlev, tv.nchild, pre, nch_pre);
    scan_exclusive(nch_pre, base_pre, nb, nullptr, scan_tmp, s);
    hipLaunchKernelGGL(k_refid, dim3(blocks(nb)), dim3(kT), 0, s, nb, pre, base_pre, tv.parent, tv.bsys,
                       tv.child_first, tv.nchild, tv.refid);

    // ---- leaves, pass 2 (input order inside each leaf), gather
    int* flags = pool.get<int>("leaf_flag", nb);
    int* fscan = pool.get<int>("leaf_scan", nb);
    // leaves before each level boundary (for the non-leaf ranges per level),
    // boxes per octant (for the M2L tiles), the leaf count and the largest
    // leaf: from the topology passes (read back with the level offsets), else
    // gathered here in one more round trip
    const int nlev = (int)meta.level_off.size();
    std::vector<int> got(nlev + 8 + 2);
    if (!topo_extra.empty()) {
        hipLaunchKernelGGL(k_leaf_flags, dim3(blocks(nb)), dim3(kT), 0, s, nb, tv.nchild, tv.seg_count, flags,
                           nullptr);
        scan_exclusive(flags, fscan, nb, nullptr, scan_tmp, s);
        int before = 0;
        for (int lv = 0; lv < nlev; ++lv) {
            got[lv] = before;  // leaves of the levels below boundary lv
            if (lv <= kMaxDepth) before += topo_extra[kTopoLeafCnt + lv];
        }
        for (int q = 0; q < 8; ++q) got[nlev + q] = topo_extra[kTopoOctCnt + q];
        got[nlev + 8] = before;
        got[nlev + 9] = topo_extra[kTopoMaxLeaf];
    } else {
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(kT), 0, s, d_total + 1, 1, 0);
        hipLaunchKernelGGL(k_leaf_flags, dim3(blocks(nb)), dim3(kT), 0, s, nb, tv.nchild, tv.seg_count, flags,
                           d_total + 1);
        scan_exclusive(flags, fscan, nb, d_total, scan_tmp, s);
        int* up = pool.get<int>("lvl_idx", nlev);
        AFMM_CHECK(hipMemcpyAsync(up, meta.level_off.data(), sizeof(int) * nlev, hipMemcpyHostToDevice, s));
        int* d_lvl_out = pool.get<int>("lvl_out", nlev + 8 + 2);
        hipLaunchKernelGGL(k_gather_at, dim3(1), dim3(64), 0, s, fscan, up, nlev, nb, d_lvl_out, d_total, 2);
        hipLaunchKernelGGL(k_octant_count, dim3(blocks(nb, kOctCountThreads * kOctCountPer)),
                           dim3(kOctCountThreads), 0, s, nb, tv.octant, d_lvl_out + nlev);
        d_oct_cnt = d_lvl_out + nlev;
        AFMM_CHECK(hipMemcpyAsync(got.data(), d_lvl_out, sizeof(int) * (nlev + 8 + 2), hipMemcpyDeviceToHost, s));
        AFMM_CHECK(hipStreamSynchronize(s));
    }
    tv.nleaves = got[nlev + 8];
    const int max_leaf = got[nlev + 9];
    tv.leaf_ids = pool.get<int>("leaf_ids", tv.nleaves);
    hipLaunchKernelGGL(k_leaf_ids, dim3(blocks(nb)), dim3(kT), 0, s, nb, tv.nchild, fscan, tv.leaf_ids);
    tv.leaf_of = pool.get<int>("leaf_of", n);
    // input order inside the leaves, and the bodies gathered into it: ranked
    // per leaf when leaves are small (one kernel), else a stable radix sort of
    // all bodies by leaf start and a gather
    tv.x = pool.get<double>("bx", n);
    tv.y = pool.get<double>("by", n);
    tv.z = pool.get<double>("bz", n);
    tv.m = pool.get<double>("bm", n);
    tv.orig = pool.get<int>("orig", n);
    if (max_leaf <= kLeafRankMax) {
        hipLaunchKernelGGL(k_leaf_order, dim3(blocks(64LL * tv.nleaves)), dim3(kT), 0, s, tv.nleaves, tv.leaf_ids,
                           tv.seg_first, tv.seg_count, perm1, d_pos, d_mass, tv.x, tv.y, tv.z, tv.m, tv.orig,
                           tv.leaf_of);
    } else {
        SortBufs lk;  // buffers of their own: the pass-1 keys in sb are read later
        for (int k = 0; k < 2; ++k) {
            lk.khi[k] = pool.get<uint64_t>(k ? "lk_khi1" : "lk_khi0", n);
            lk.klo[k] = pool.get<uint32_t>(k ? "lk_klo1" : "lk_klo0", n);
            lk.val[k] = pool.get<int>(k ? "lk_val1" : "lk_val0", n);
        }
        lk.hist = sb.hist;
        lk.scan_tmp = sb.scan_tmp;
        hipLaunchKernelGGL(k_leaf_keys, dim3(blocks(64LL * tv.nleaves)), dim3(kT), 0, s, tv.nleaves, tv.leaf_ids,
                           tv.seg_first, tv.seg_count, perm1, lk.khi[0], lk.klo[0], lk.val[0], tv.leaf_of);
        int bits = 1;
        while ((1LL << bits) <= n) ++bits;
        const int* leaf_perm = lk.val[radix_sort(lk, n, 1, (bits + 8) / 9, s)];
        hipLaunchKernelGGL(k_gather, dim3(blocks(n)), dim3(kT), 0, s, n, leaf_perm, d_pos, d_mass, tv.x, tv.y,
                           tv.z, tv.m, tv.orig);
    }

    // ---- geometry
    tv.cfp = pool.get<double>("cfp", 3 * (size_t)nb);
    tv.cdd = pool.get<double>("cdd", 6 * (size_t)nb);
    tv.half = pool.get<double>("half", nb);
    tv.invh = pool.get<double>("invh", nb);
    tv.ic = pool.get<int>("ic", 3 * (size_t)nb);
    hipLaunchKernelGGL(k_geometry, dim3(blocks(nb)), dim3(kT), 0, s, nb, d_roots, tv.bsys, tv.seg_first, tv.level,
                       khi, klo, tv.cfp, tv.cdd, tv.half, tv.invh, tv.ic);

    // ---- non-leaf boxes per level (what the upward pass reads besides the
    // bodies, leaves and geometry)
    tv.nonleaf_ids = pool.get<int>("nonleaf_ids", nb);
    hipLaunchKernelGGL(k_nonleaf_ids, dim3(blocks(nb)), dim3(kT), 0, s, nb, tv.nchild, fscan, tv.nonleaf_ids);
    {
        const std::vector<int>& bounds = meta.level_off;
        meta.nonleaf_off.resize(nlev);
        for (int l = 0; l < nlev; ++l)
            meta.nonleaf_off[l] = bounds[l] - (bounds[l] < nb ? got[l] : tv.nleaves);
    }

    if (on_core) on_core(tv, meta);

    // ---- (level, octant) order and the M2L target order and tiles (after the
    // hook: the upward pass needs none of them)
    {
        SortBufs lb;
        for (int k = 0; k < 2; ++k) {
            lb.khi[k] = pool.get<uint64_t>(k ? "lo_khi1" : "lo_khi0", nb);
            lb.klo[k] = pool.get<uint32_t>(k ? "lo_klo1" : "lo_klo0", nb);
            lb.val[k] = pool.get<int>(k ? "lo_val1" : "lo_val0", nb);
        }
        lb.hist = pool.get<int>("lo_hist", 512 * (size_t)sort_blocks(nb));
        lb.scan_tmp = pool.get<int>("lo_scan", scan_tmp_ints(512 * sort_blocks(nb) + nb) + 16);
        hipLaunchKernelGGL(k_lo_keys, dim3(blocks(nb)), dim3(kT), 0, s, nb, tv.level, tv.octant, lb.khi[0],
                           lb.klo[0], lb.val[0]);
        // (the sort's value buffer is the order: nothing sorts with lb again)
        tv.lo_order = lb.val[radix_sort(lb, nb, 1, 1, s)];
        tv.m2l_order = pool.get<int>("m2l_order", nb);
        if (nb > 1)  // (a single box: the sort ran no pass and left no histogram)
            hipLaunchKernelGGL(k_m2l_from_lo, dim3(blocks(nb)), dim3(kT), 0, s, nb, tv.lo_order, tv.level,
                               tv.octant, lb.hist, sort_blocks(nb), tv.m2l_order);
        else
            hipLaunchKernelGGL(k_fill, dim3(1), dim3(kT), 0, s, tv.m2l_order, nb, 0);
        // M2L tiles (tile starts built on the device from the same counts)
        int tot = 0;
        tv.m2l_ntiles256 = tv.m2l_ntiles128 = 0;
        const int* per_octant = got.data() + nlev;
        const int cus = device_cus();
        const int t256 = m2l_tile_size(per_octant, 256, cus), t128 = m2l_tile_size(per_octant, 128, cus);
        tv.m2l_t256 = t256;
        for (int q = 0; q < 8; ++q) {
            const int c = per_octant[q];
            if (c <= 0) continue;
            tot += c;
            tv.m2l_ntiles256 += m2l_tiles_of(c, t256);
            tv.m2l_ntiles128 += m2l_tiles_of(c, t128);
        }
        if (tot != nb) throw std::runtime_error("octant counts do not cover the boxes");
        tv.m2l_tiles256 = pool.get<int>("m2l_tiles256", tv.m2l_ntiles256 + 1);
        tv.m2l_tiles128 = pool.get<int>("m2l_tiles128", tv.m2l_ntiles128 + 1);
        hipLaunchKernelGGL(k_m2l_tile_plan, dim3(1), dim3(kT), 0, s, d_oct_cnt, t256, tv.m2l_ntiles256,
                           tv.m2l_tiles256);
        hipLaunchKernelGGL(k_m2l_tile_plan, dim3(1), dim3(kT), 0, s, d_oct_cnt, t128, tv.m2l_ntiles128,
                           tv.m2l_tiles128);
    }

    // ---- colleagues, V, M2L source table
    tv.coll = pool.get<int>("coll", kCollSlots * (size_t)nb);
    tv.ncoll = pool.get<int>("ncoll", nb);
    tv.vl = pool.get<int>("vl", kVSlots * (size_t)nb);
    tv.nv = pool.get<int>("nv", nb);
    int* overflow = pool.get<int>("overflow", 1);
    int* aset = pool.get<int>("aset", kCollSlots * (size_t)nb);
    int* acnt = pool.get<int>("acnt", nb);
    int* xset = pool.get<int>("xset", kCollSlots * (size_t)nb);
    int* xcnt = pool.get<int>("xcnt", nb + 1);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(kT), 0, s, overflow, 1, 0);
    for (int lv = 0; lv <= tv.max_level; ++lv) {
        const int off = meta.level_off[lv], cnt = meta.level_off[lv + 1] - off;
        hipLaunchKernelGGL(k_coll_v, dim3(cnt), dim3(64), 0, s, off, d_roots, tv.bsys, tv.parent, tv.child_first,
                           tv.nchild, tv.child_mask, tv.cfp, tv.half, tv.ic, tv.refid, tv.coll, tv.ncoll, tv.vl,
                           tv.nv, aset, acnt, xset, xcnt, overflow);
    }
    // (the overflow flag is read back with the list sizes below; list counts
    // are capped at their slots, so the kernels in between stay in bounds)
    // (the list stream waits for the colleague lists before on_vlists enqueues
    // more work on s, which the lists do not need)
    if (ls && ls != s) {
        AFMM_CHECK(hipEventRecord(ls_ready, s));
        AFMM_CHECK(hipStreamWaitEvent(ls, ls_ready, 0));
    }
    if (on_vlists) on_vlists(tv, meta);

    // ---- U and W per leaf, X per box (count, scan, one size readback, fill, order)
    if (ls && ls != s) s = ls;
    const int nl = tv.nleaves;
    int* ucnt = pool.get<int>("ucnt", nl + 1);
    int* wcnt = pool.get<int>("wcnt", nl + 1);
    int* ucur = pool.get<int>("ucur", nl + 1);
    int* wcur = pool.get<int>("wcur", nl + 1);
    int* usrc = pool.get<int>("usrc", nl + 1);
    tv.u_off = pool.get<int>("u_off", nl + 1);
    tv.w_off = pool.get<int>("w_off", nl + 1);
    tv.x_off = pool.get<int>("x_off", nb + 1);
    hipLaunchKernelGGL(k_zero4, dim3(blocks(nl)), dim3(kT), 0, s, nl, ucnt, wcnt, ucur, wcur);
    hipLaunchKernelGGL(k_list_count, dim3(blocks((long long)kCollSlots * nb)), dim3(kT), 0, s, nb, tv.nchild, tv.level, fscan, aset, acnt,
                       xset, xcnt, ucnt, wcnt);
    scan_exclusive(ucnt, tv.u_off, nl, tv.u_off + nl, scan_tmp, s);
    scan_exclusive(wcnt, tv.w_off, nl, tv.w_off + nl, scan_tmp + scan_tmp_ints(scan_cap), s);
    scan_exclusive(xcnt, tv.x_off, nb, tv.x_off + nb, scan_tmp, s);
    {
        int tot[4];
        int* d_tot = pool.get<int>("list_sizes", 4);
        tv.list_totals = pool.get<unsigned long long>("list_totals", 4);
        hipLaunchKernelGGL(k_list_sizes, dim3(1), dim3(64), 0, s, tv.u_off + nl, tv.w_off + nl, tv.x_off + nb, overflow,
                           d_tot, tv.list_totals);
        AFMM_CHECK(hipMemcpyAsync(tot, d_tot, sizeof(tot), hipMemcpyDeviceToHost, s));
        AFMM_CHECK(hipStreamSynchronize(s));
        if (tot[3] != 0) throw std::runtime_error("interaction-list slot overflow");
        tv.n_u = tot[0];
        tv.n_w = tot[1];
        tv.n_x = tot[2];
    }
    tv.u_idx = pool.get<int>("u_idx", tv.n_u);
    tv.w_idx = pool.get<int>("w_idx", tv.n_w);
    tv.x_idx = pool.get<int>("x_idx", tv.n_x);
    hipLaunchKernelGGL(k_list_fill, dim3(blocks((long long)kCollSlots * nb)), dim3(kT), 0, s, nb, tv.nchild, tv.level, fscan, aset, acnt,
                       xset, xcnt, tv.u_off, tv.w_off, tv.x_off, ucur, wcur, tv.u_idx, tv.w_idx, tv.x_idx);
    hipLaunchKernelGGL(k_list_sort, dim3(nl), dim3(64), 0, s, tv.u_off, tv.w_off, tv.seg_first, tv.seg_count,
                       tv.u_idx, tv.w_idx, usrc);
    hipLaunchKernelGGL(k_list_totals, dim3(blocks(std::max(nb, nl), kTotThreads)), dim3(kTotThreads), 0, s, nb, nl,
                       tv.ncoll, tv.nv, tv.leaf_ids, tv.seg_count, usrc, tv.list_totals);
    AFMM_CHECK(hipGetLastError());
}

}  // namespace afmm
