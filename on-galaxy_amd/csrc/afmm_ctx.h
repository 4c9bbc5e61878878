// ============================================================================
// afmm_ctx.h -- the solver context behind the opaque afmm_ctx of afmm.h and
// the evaluation entry shared by the C ABI, the resident integrator and the
// request server.
// ============================================================================
#pragma once

#include <string>

#include "../../include/afmm.h"
#include "afmm_internal.h"

struct afmm_ctx {
    afmm_params prm{};
    hipStream_t stream = nullptr;
    afmm::Pool pool;
    std::string err;
    // tables
    double *d_m2m_sw = nullptr, *d_l2l_sw = nullptr, *d_m2l_sw = nullptr;
    float2* d_m2l_hl = nullptr;  // split FP32 M2L operators (AFMM_PRECISION_MIXED only)
    uint4* d_m2l_bf = nullptr;   // three-piece bf16 M2L operators (AFMM_PRECISION_MIXED_BF16 only)
    // last evaluation
    bool have_tree = false;
    afmm::TreeView tv{};
    afmm::TreeMeta meta;
    afmm::RootInfo root{};  // of system 0
    int key_levels_hint = -1;  // key depth of the last tree (the next build's guess)
    int tree_depth_hint = -1;  // deepest level of the last tree (the levels the next sort orders)
    afmm::Shapes shapes{};  // kernel shapes (afmm_set_shapes), defaults resolved at create
    afmm::HostStager stage;  // pinned staging of host arrays (afmm_eval, afmm_eval_batch)
    int nsys = 1;
    afmm_stats stats{};
    hipEvent_t ev[10] = {};
    // second stream (the process's shared one, taken on first use) for the
    // upward pass beside the list build and the P2P near field beside the
    // far-field passes; off in graph mode and for the request server's
    // workers, which already run several evaluations side by side
    bool split_p2p = true;
    int overlap_mode = AFMM_OVERLAP_AUTO;  // afmm_set_overlap
    hipStream_t side = nullptr;
    hipEvent_t side_fork = nullptr, side_join = nullptr, up_fork = nullptr, up_done = nullptr;
    // the U, W and X lists on the second stream beside M2L (overlap on, graph off)
    hipEvent_t list_fork = nullptr, lists_done = nullptr, m2l_done = nullptr, x_done = nullptr, w_done = nullptr;
    // FMM passes replayed as a HIP graph (afmm_set_graph): the executable
    // graph of the last evaluation, updated in place when the next capture has
    // the same shape, re-instantiated otherwise
    bool use_graph = false;
    hipStream_t graph_side = nullptr;  // the captured P2P branch's stream (graph mode with the overlap on)
    hipGraphExec_t gexec = nullptr;
    int64_t graph_instantiations = 0, graph_updates = 0;
};

namespace afmm {

inline int fail(afmm_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

template <class F>
inline int guarded(afmm_ctx* c, F&& f) {
    try {
        return f();
    } catch (const HipError& e) {
        const bool oom = e.code == hipErrorOutOfMemory || e.code == hipErrorMemoryAllocation;
        return fail(c, oom ? AFMM_ERR_NO_MEMORY : AFMM_ERR_DEVICE, e.what());
    } catch (const std::bad_alloc&) {
        return fail(c, AFMM_ERR_NO_MEMORY, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(c, AFMM_ERR_DEVICE, e.what());
    }
}

// Forest evaluation of nsys systems on device arrays (concatenated pos AoS
// n x 3 and mass; system q = bodies [sys_off[q], sys_off[q+1]), host offsets)
// into acc (and phi when non-null) with per-system G and eps (host arrays).
// Synchronises the context's stream; records the statistics of the last tree.
int eval_forest(afmm_ctx* c, int nsys, const int* sys_off, const double* G, const double* eps, const double* pos,
                const double* mass, double* acc, double* phi);

}  // namespace afmm
