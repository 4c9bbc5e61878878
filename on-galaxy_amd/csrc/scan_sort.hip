// ============================================================================
// scan_sort.hip -- device-wide exclusive scan and a stable LSD radix sort with
// 9-bit digits (three octree levels per pass), written for wave64.
//
// Sort pass = three kernels: per-tile digit histograms (LDS atomics on
// integer counters), one exclusive scan over the digit-major histogram, and
// a scatter that ranks every item stably inside its tile. In-tile ranks come
// from wave ballots over the nine digit bits (lanes with the same digit form
// a match mask; rank = popcount of lower lanes), then a per-digit prefix over
// the four waves of the block. Everything is integer arithmetic: the result
// is deterministic and identical to a sequential stable counting sort.
// ============================================================================
#include <algorithm>
#include <stdexcept>

#include "afmm_internal.h"
#include "block_scan.h"

namespace afmm {
namespace {

// rng (optional, device): the item count is rng[1] - rng[0] instead of n, and
// blocks past it only write zero tile sums.
__device__ __forceinline__ int scan_count(int n, const int* rng) {
    return rng ? max(0, min(n, rng[1] - rng[0])) : n;  // n bounds the range (buffer size)
}

__global__ void __launch_bounds__(kScanThreads) k_tile_sums(const int* in, int n, int* sums, const int* rng) {
    __shared__ int lds[kScanThreads];
    n = scan_count(n, rng);
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += in[base + k];
    int total;
    block_exclusive_scan(s, lds, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Each block finds its tile's offset by summing the raw tile sums before it
// (a few hundred ints at most for the sizes scanned here), then scans its own
// tile; the block of the last tile writes the total (with rng: rng[1] + total,
// the end offset of the next range). Two launches per scan instead of three.
__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const int* in, int n, const int* sums,
                                                             int* out, const int* rng, int* total_out) {
    __shared__ int lds[kScanThreads];
    n = scan_count(n, rng);
    const int last = n > 0 ? (n - 1) / kScanTile : 0;
    if ((int)blockIdx.x > last) return;
    const int tile_base = block_prefix_of_tiles(sums, blockIdx.x, lds);
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int total;
    int run = block_exclusive_scan(s, lds, total) + tile_base;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if ((int)blockIdx.x == last && threadIdx.x == 0 && total_out)
        *total_out = tile_base + total + (rng ? rng[1] : 0);
}

// ---- radix sort -----------------------------------------------------------
constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / kWave;
constexpr int kSortRounds = 8;
constexpr int kSortTile = kSortThreads * kSortRounds;  // 2048 items per block
constexpr int kBuckets = 512;

// digit g: 0 = klo bits [0, 9); 1..7 = khi bits [9(g-1), 9g); 8, 9 = klo bits
// [9, 18), [18, 27) (the system index of a forest, above every path digit).
__device__ __forceinline__ int digit_of(uint64_t khi, uint32_t klo, int g) {
    if (g == 0) return int(klo & 511u);
    if (g >= 8) return int((klo >> (9 * (g - 7))) & 511u);
    return int((khi >> (9 * (g - 1))) & 511u);
}

// LO = false: the low key words are all zero (one system, at most 21 levels)
// and are neither read nor moved; digits then come from khi alone (g in 1..7).
template <bool LO>
__global__ void __launch_bounds__(kSortThreads) k_sort_hist(const uint64_t* khi, const uint32_t* klo,
                                                            int n, int g, int* hist, int nblocks) {
    __shared__ int h[kBuckets];
    for (int d = threadIdx.x; d < kBuckets; d += kSortThreads) h[d] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortRounds; ++r) {
        const int i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[digit_of(khi[i], LO ? klo[i] : 0u, g)], 1);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kBuckets; d += kSortThreads) hist[d * nblocks + blockIdx.x] = h[d];
}

// Stable scatter of one tile: the tile's items are first ranked into LDS in
// (digit, input order) order, then written out in that order, so that
// consecutive threads write consecutive addresses of each digit's run (a
// direct scatter sends a wave's 64 writes to up to 64 runs: measured 2.3x
// slower on passes with many distinct digits). Local digit starts come from
// the tile's own histogram, global ones from the scanned histogram offs.
template <bool LO>
__global__ void __launch_bounds__(kSortThreads) k_sort_scatter(const uint64_t* khi, const uint32_t* klo,
                                                               const int* val, int n, int g,
                                                               const int* offs, int nblocks,
                                                               uint64_t* okhi, uint32_t* oklo, int* oval) {
    __shared__ uint64_t skh[kSortTile];
    __shared__ uint32_t skl[LO ? kSortTile : 1];
    __shared__ int sval[kSortTile];
    __shared__ int run[kBuckets], lstart[kBuckets], gstart[kBuckets];
    __shared__ int cnt[kSortWaves][kBuckets];
    __shared__ int wsum[kSortWaves];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int base = blockIdx.x * kSortTile;
    const int m = min(kSortTile, n - base);  // items in this tile (> 0)
    for (int d = threadIdx.x; d < kBuckets; d += kSortThreads) {
        run[d] = 0;
        gstart[d] = offs[d * nblocks + blockIdx.x];
        for (int w = 0; w < kSortWaves; ++w) cnt[w][d] = 0;
    }
    __syncthreads();
    // the tile's digit histogram, then its exclusive scan (two digits per thread)
    for (int r = 0; r < kSortRounds; ++r) {
        const int i = r * kSortThreads + threadIdx.x;
        if (i < m) atomicAdd(&run[digit_of(khi[base + i], LO ? klo[base + i] : 0u, g)], 1);
    }
    __syncthreads();
    {
        static_assert(kBuckets == 2 * kSortThreads, "two digits per thread");
        const int d0 = 2 * threadIdx.x;
        const int c0 = run[d0], c1 = run[d0 + 1];
        int incl = c0 + c1;
#pragma unroll
        for (int k = 1; k < kWave; k <<= 1) {
            const int t = __shfl_up(incl, k, kWave);
            if (lane >= k) incl += t;
        }
        if (lane == kWave - 1) wsum[wave] = incl;
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        const int ex = before + incl - c0 - c1;
        lstart[d0] = ex;
        lstart[d0 + 1] = ex + c0;
        run[d0] = ex;
        run[d0 + 1] = ex + c0;
    }
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < kSortRounds; ++r) {
        const int i = r * kSortThreads + threadIdx.x;
        const bool act = i < m;
        uint64_t kh = 0;
        uint32_t kl = 0;
        int v = 0, d = 0;
        if (act) {
            kh = khi[base + i];
            if constexpr (LO) kl = klo[base + i];
            v = val[base + i];
            d = digit_of(kh, kl, g);
        }
        uint64_t match = __ballot(act);
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            const bool bit = (d >> b) & 1;
            const uint64_t vote = __ballot(bit);
            match &= bit ? vote : ~vote;
        }
        const int rank = __popcll(match & lt);
        if (act && rank == 0) cnt[wave][d] = __popcll(match);
        __syncthreads();
        if (act) {
            int pos = run[d] + rank;
            for (int w = 0; w < wave; ++w) pos += cnt[w][d];
            skh[pos] = kh;
            if constexpr (LO) skl[pos] = kl;
            sval[pos] = v;
        }
        __syncthreads();
        for (int dd = threadIdx.x; dd < kBuckets; dd += kSortThreads) {
            int s2 = 0;
            for (int w = 0; w < kSortWaves; ++w) {
                s2 += cnt[w][dd];
                cnt[w][dd] = 0;
            }
            run[dd] += s2;
        }
        __syncthreads();
    }
    // the staged tile in digit order: item j goes to its digit's global run
    for (int j = threadIdx.x; j < m; j += kSortThreads) {
        const uint64_t kh = skh[j];
        const uint32_t kl = LO ? skl[j] : 0u;
        const int d = digit_of(kh, kl, g);
        const int pos = gstart[d] + (j - lstart[d]);
        okhi[pos] = kh;
        if constexpr (LO) oklo[pos] = kl;
        oval[pos] = sval[j];
    }
}

}  // namespace

int scan_tmp_ints(int n) { return (n + kScanTile - 1) / kScanTile + 1; }

void scan_exclusive(const int* d_in, int* d_out, int n, int* d_total, int* d_tmp, hipStream_t s) {
    n = std::max(n, 0);
    const int tiles = std::max(1, (n + kScanTile - 1) / kScanTile);
    if (n > 0) hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(kScanThreads), 0, s, d_in, n, d_tmp, nullptr);
    hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(kScanThreads), 0, s, d_in, n, d_tmp, d_out, nullptr, d_total);
}

void scan_exclusive_range(const int* d_in, int* d_out, const int* d_rng, int nmax, int* d_end, int* d_tmp,
                          hipStream_t s) {
    const int tiles = std::max(1, (nmax + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(k_tile_sums, dim3(tiles), dim3(kScanThreads), 0, s, d_in, nmax, d_tmp, d_rng);
    hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(kScanThreads), 0, s, d_in, nmax, d_tmp, d_out, d_rng, d_end);
}

int sort_blocks(int n) { return (n + kSortTile - 1) / kSortTile; }

int radix_sort(SortBufs& b, int n, int first_digit, int ndigits, hipStream_t s, int start, bool lo) {
    int cur = start;
    if (n <= 1) return cur;
    const int nblocks = sort_blocks(n);
    if (!lo && (first_digit < 1 || first_digit + ndigits > 8))
        throw std::runtime_error("radix_sort: digits outside khi without the low key words");
    for (int k = 0; k < ndigits; ++k) {
        const int g = first_digit + k;
        if (lo)
            hipLaunchKernelGGL(k_sort_hist<true>, dim3(nblocks), dim3(kSortThreads), 0, s, b.khi[cur], b.klo[cur],
                               n, g, b.hist, nblocks);
        else
            hipLaunchKernelGGL(k_sort_hist<false>, dim3(nblocks), dim3(kSortThreads), 0, s, b.khi[cur], nullptr,
                               n, g, b.hist, nblocks);
        scan_exclusive(b.hist, b.hist, kBuckets * nblocks, nullptr, b.scan_tmp, s);
        if (lo)
            hipLaunchKernelGGL(k_sort_scatter<true>, dim3(nblocks), dim3(kSortThreads), 0, s, b.khi[cur], b.klo[cur],
                               b.val[cur], n, g, b.hist, nblocks, b.khi[1 - cur], b.klo[1 - cur], b.val[1 - cur]);
        else
            hipLaunchKernelGGL(k_sort_scatter<false>, dim3(nblocks), dim3(kSortThreads), 0, s, b.khi[cur], nullptr,
                               b.val[cur], n, g, b.hist, nblocks, b.khi[1 - cur], nullptr, b.val[1 - cur]);
        cur = 1 - cur;
    }
    return cur;
}

}  // namespace afmm
