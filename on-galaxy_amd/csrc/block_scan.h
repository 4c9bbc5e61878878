// ============================================================================
// block_scan.h -- the block-wide exclusive scan shared by the device-wide scan
// (scan_sort.hip) and the fused scan of the level-by-level tree build
// (tree.hip): 256 threads of 8 items each per 2048-item tile.
// ============================================================================
#pragma once

#include "afmm_internal.h"

namespace afmm {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048

// Exclusive scan of one int per thread over the block: wave inclusive scans
// (shuffles), then the wave totals through LDS (lds: at least kScanThreads / 64
// ints). Integer sums: the result does not depend on the order.
__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int& total) {
    constexpr int kW = kScanThreads / kWave;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
    int incl = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int t = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += t;
    }
    __syncthreads();  // lds free from a previous call
    if (lane == kWave - 1) lds[w] = incl;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int k = 0; k < kW; ++k) {
        const int t = lds[k];
        before += k < w ? t : 0;
        total += t;
    }
    return before + incl - v;
}

// Sum of sums[0 .. upto) over the block (every thread gets it): the offset of
// tile `upto` when sums holds raw per-tile totals.
__device__ __forceinline__ int block_prefix_of_tiles(const int* sums, int upto, int* lds) {
    int part = 0;
    for (int k = threadIdx.x; k < upto; k += kScanThreads) part += sums[k];
    int total;
    block_exclusive_scan(part, lds, total);
    return total;
}

}  // namespace afmm
