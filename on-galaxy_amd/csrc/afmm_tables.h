// ============================================================================
// afmm_tables.h -- level-independent translation matrices in the scaled real
// basis of afmm_math.h, built on the host in long double.
//
//   m2m[o]  child (octant o) -> parent   (FmmExpansions3D.hpp:242-265)
//   l2l[o]  parent -> child (octant o)   (FmmExpansions3D.hpp:307-330)
//   m2l[k]  source -> target, k = offset_index(target - source)
//                                        (FmmExpansions3D.hpp:273-299)
// Each matrix is S x S (S = (p+1)^2) stored column-major: entry (row r, col c)
// at [c*S + r], so out[r] = sum_c T[c*S + r] in[c]. Offsets with all
// components in [-1, 1] are adjacent pairs and are left zero.
// ============================================================================
#pragma once

#include <cstdint>
#include <vector>

namespace afmm {

struct Tables {
    int p = 0, S = 0;
    std::vector<double> m2m;  // 8 * S * S
    std::vector<double> l2l;  // 8 * S * S
    std::vector<double> m2l;  // 343 * S * S
};

// Builds (and caches per p for the life of the process) the tables of order p.
const Tables& tables_for_order(int p);

// Re-lays S x S column-major matrices for v_mfma_f64_16x16x4_f64 A operands:
// [matrix][k chunk c][row block rb][k step kk within chunk][lane], value
// T(row = 16 rb + lane%16, col = 4 (c*KC + kk) + lane/16), zero outside
// S x S (chunking of afmm_math.h m2l_*). A chunk is contiguous and one
// fragment is one 512-byte run of 64 lanes.
std::vector<double> mfma_a_layout(const std::vector<double>& colmajor, int S, int nmat);
// M2L operators in the FP64 M2L kernel's chunks: MFMA fragments plus VALU tail rows.
std::vector<double> m2l_a_layout(const std::vector<double>& colmajor, int S, int nmat);

// The same matrices for v_mfma_f32_16x16x32_bf16, each entry split into three
// bf16 pieces (x = hi + mid + lo to about 2^-24 relative; bf16 round to nearest
// even of the running remainder): [matrix][k chunk c][row block rb][k block kk]
// [piece][lane][8], value T(row = 16 rb + lane%16, col = 32 (c*KC + kk) +
// 8 (lane/16) + e), zero outside S x S (chunking of afmm_math.h m2l_bf_*).
std::vector<uint16_t> mfma_bf16_layout(const std::vector<double>& colmajor, int S, int nmat);
// Three-piece bf16 split of x (bits), as above.
void bf16_split3(double x, uint16_t out[3]);

}  // namespace afmm
