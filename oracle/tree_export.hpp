// ============================================================================
// tree_export.hpp -- TEST INFRASTRUCTURE ONLY. Copies an octree with
// reference field names (AdaptiveOctree.hpp:31-47) into the afmm_tree export
// layout, so the oracle, the compiled reference and the HIP library can be
// compared field by field. Works for any tree type exposing box[] (parent,
// child[8], level, cx, cy, cz, half, leaf, pfirst, pcount), the permuted SoA,
// origIndex, leaves and the five list vectors.
// ============================================================================
#pragma once

#include <cstdint>
#include <vector>

#include "../include/afmm.h"

namespace oracle_export {

inline int64_t total(const std::vector<std::vector<int>>& lists) {
    int64_t s = 0;
    for (const auto& l : lists) s += (int64_t)l.size();
    return s;
}

inline void csr(const std::vector<std::vector<int>>& lists, int32_t* off, int32_t* idx) {
    int64_t k = 0;
    if (off) off[0] = 0;
    for (std::size_t b = 0; b < lists.size(); ++b) {
        for (int v : lists[b]) {
            if (idx) idx[k] = v;
            ++k;
        }
        if (off) off[b + 1] = (int32_t)k;
    }
}

// sizes: n_bodies, n_boxes, n_leaves, n_colleague, n_u, n_v, n_w, n_x
template <class BoxVec, class ListVec>
void sizes(const BoxVec& box, std::size_t nBodies, std::size_t nLeaves, const ListVec& coll,
           const ListVec& U, const ListVec& V, const ListVec& W, const ListVec& X, int64_t* out) {
    out[0] = (int64_t)nBodies;
    out[1] = (int64_t)box.size();
    out[2] = (int64_t)nLeaves;
    out[3] = total(coll);
    out[4] = total(U);
    out[5] = total(V);
    out[6] = total(W);
    out[7] = total(X);
}

template <class BoxVec>
void boxes(const BoxVec& box, afmm_tree* t) {
    for (std::size_t b = 0; b < box.size(); ++b) {
        const auto& x = box[b];
        if (t->parent) t->parent[b] = x.parent;
        if (t->child)
            for (int o = 0; o < 8; ++o) t->child[8 * b + o] = x.child[o];
        if (t->level) t->level[b] = x.level;
        if (t->center) {
            t->center[3 * b] = x.cx;
            t->center[3 * b + 1] = x.cy;
            t->center[3 * b + 2] = x.cz;
        }
        if (t->half) t->half[b] = x.half;
        if (t->leaf) t->leaf[b] = x.leaf ? 1 : 0;
        if (t->pfirst) t->pfirst[b] = x.pfirst;
        if (t->pcount) t->pcount[b] = x.pcount;
    }
}

inline void bodies(const std::vector<double>& px, const std::vector<double>& py,
                   const std::vector<double>& pz, const std::vector<double>& m,
                   const std::vector<int>& orig, afmm_tree* t) {
    for (std::size_t k = 0; k < px.size(); ++k) {
        if (t->px) t->px[k] = px[k];
        if (t->py) t->py[k] = py[k];
        if (t->pz) t->pz[k] = pz[k];
        if (t->mass) t->mass[k] = m[k];
        if (t->orig_index) t->orig_index[k] = orig[k];
    }
}

}  // namespace oracle_export
