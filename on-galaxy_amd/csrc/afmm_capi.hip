// ============================================================================
// afmm_capi.hip -- the C ABI of include/afmm.h.
//
// A context owns a HIP stream, grow-only device pools (no per-call
// hipMalloc/hipFree once sizes stabilise) and the order-p translation tables.
// afmm_eval is the reference's drop-in semantics (host arrays in and out,
// AdaptiveFMM3D::computeAccelerations, AdaptiveFMM3D.hpp:29-110);
// afmm_eval_device is the same evaluation on device-resident arrays.
// No function exits the process: HIP errors become AFMM_ERR_DEVICE with the
// message kept for afmm_last_error.
// ============================================================================
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/afmm.h"
#include "afmm_ctx.h"
#include "afmm_internal.h"
#include "afmm_math.h"
#include "afmm_tables.h"

using namespace afmm;

namespace {

int check_params(const afmm_params* p) {
    if (!p) return AFMM_ERR_ARG;
    if (!(std::isfinite(p->G)) || !(std::isfinite(p->eps)) || p->eps < 0) return AFMM_ERR_ARG;
    if (p->p < 1 || p->p > kMaxOrder) return AFMM_ERR_ORDER;
    if (p->leaf_cap < 1) return AFMM_ERR_LEAF_CAP;
    if (p->precision < AFMM_PRECISION_FP64 || p->precision > AFMM_PRECISION_P2P_FP32)
        return AFMM_ERR_PRECISION;
    return AFMM_OK;
}

// The second stream for P2P, one per device for the whole process (created on
// first use, never destroyed): contexts share it through their own fork and
// join events, so extra contexts add no streams (each stream beyond the few
// hardware queues shares a queue with another and can serialise behind it).
hipStream_t shared_side_stream() {
    static std::mutex mu;
    static std::map<int, hipStream_t> streams;
    int dev = 0;
    AFMM_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    hipStream_t& st = streams[dev];
    if (!st) AFMM_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    return st;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

}  // namespace

// The evaluation of nsys systems (concatenated device arrays: pos AoS n x 3,
// mass n; system q = bodies [sys_off[q], sys_off[q+1])) into acc (AoS, input
// order) and optionally phi. G and eps per system (host arrays).
// FP64 at p = 9-13: the most bodies for which the automatic schedule runs the
// list, P2L, M2P and P2P passes beside M2L.
constexpr int64_t kBesideMaxBodiesMidOrders = 500000;

int afmm::eval_forest(afmm_ctx* c, int nsys, const int* sys_off, const double* G, const double* eps, const double* pos,
                const double* mass, double* acc, double* phi) {
    c->have_tree = false;
    const int64_t n = sys_off[nsys];
    if (n == 0) return AFMM_OK;
    hipStream_t s = c->stream;
    // per-system constants (uploaded before the build's first synchronisation)
    std::vector<double> h_ge(2 * (size_t)nsys);
    for (int q = 0; q < nsys; ++q) h_ge[q] = G[q], h_ge[nsys + q] = eps[q] * eps[q];
    double* d_ge = c->pool.get<double>("sys_G_eps2", 2 * (size_t)nsys);
    AFMM_CHECK(hipEventRecord(c->ev[0], s));
    AFMM_CHECK(hipMemcpyAsync(d_ge, h_ge.data(), sizeof(double) * 2 * nsys, hipMemcpyHostToDevice, s));
    TreeView tv{};
    TreeMeta meta;
    std::vector<RootInfo> roots;
    // Outside graph mode the upward pass runs on the second stream while the
    // interaction lists are built (both leave most of the GPU idle).
    // (also in graph mode: the graph then holds the passes from M2L on)
    const bool split = c->split_p2p;
    if (split && !c->side) {
        AFMM_CHECK(hipEventCreateWithFlags(&c->side_fork, hipEventDisableTiming));
        AFMM_CHECK(hipEventCreateWithFlags(&c->side_join, hipEventDisableTiming));
        AFMM_CHECK(hipEventCreateWithFlags(&c->up_fork, hipEventDisableTiming));
        AFMM_CHECK(hipEventCreateWithFlags(&c->up_done, hipEventDisableTiming));
        c->side = shared_side_stream();
    }
    // Beside M2L (outside graph mode, overlap on): M2L is enqueued on s right
    // after its preparation (behind the V lists), and the U, W and X lists,
    // P2L, M2P and P2P run on the second stream beside it, after the upward
    // pass: their VALU work shares the SIMDs with the MFMA-bound M2L, and the
    // lists' size readback synchronises the second stream only. The
    // automatic choice, from the merger at 16k-1M bodies: beside for the bf16
    // (1M, p = 8: 4.34 -> 4.13 ms) and split-FP32 (7.75 -> 7.54 ms) forms, and
    // for FP64 up to p = 8 (1M: 5.79 -> 5.67 ms at p = 8, 2.51 -> 2.45 ms at
    // p = 4; 16k: 1.48 -> 1.26 ms at p = 8) and from p = 14 (16-wave
    // row-split M2L: 44.8 -> 41.9 ms at p = 14, 78.8 -> 73.8 ms at p = 16);
    // FP64 at p = 9-13 beside up to 500k bodies (p = 12: 7.26 -> 7.00 ms at
    // 16k, 13.7 -> 13.0 ms at 400k) and in order above, where the longer M2L
    // loses more beside the VALU passes than they save (1M, p = 12: 23.9 ms
    // against 27.5 ms beside).
    const bool bf16_m2l = m2l_form(c->prm.precision) == kPrecMixedBf16;
    const bool auto_beside = m2l_form(c->prm.precision) != kPrecFP64 || c->prm.p <= 8 || c->prm.p >= 14 ||
                             n <= kBesideMaxBodiesMidOrders;
    const bool lists_beside = split && !c->use_graph &&
                              (c->overlap_mode == AFMM_OVERLAP_BESIDE_M2L ||
                               (c->overlap_mode == AFMM_OVERLAP_AUTO && auto_beside));
    if (lists_beside && !c->list_fork) {
        AFMM_CHECK(hipEventCreateWithFlags(&c->list_fork, hipEventDisableTiming));
        AFMM_CHECK(hipEventCreateWithFlags(&c->lists_done, hipEventDisableTiming));
        AFMM_CHECK(hipEventCreateWithFlags(&c->m2l_done, hipEventDisableTiming));
        AFMM_CHECK(hipEventCreateWithFlags(&c->x_done, hipEventDisableTiming));
        AFMM_CHECK(hipEventCreateWithFlags(&c->w_done, hipEventDisableTiming));
    }
    tv.sys_G = d_ge;
    tv.sys_eps2 = d_ge + nsys;
    FmmTables tab{c->d_m2m_sw, c->d_l2l_sw, c->d_m2l_sw, c->d_m2l_hl, c->d_m2l_bf};
    bool up_launched = false;
    auto on_core = [&](const TreeView& t, const TreeMeta& m) {
        AFMM_CHECK(hipEventRecord(c->up_fork, s));
        AFMM_CHECK(hipStreamWaitEvent(c->side, c->up_fork, 0));
        run_upward(c->pool, c->side, t, m, c->prm.p, tab, c->ev[2], c->ev[3]);
        AFMM_CHECK(hipEventRecord(c->up_done, c->side));
        up_launched = true;
    };
    // ... and the M2L preparation is enqueued right after the V lists, ahead of
    // the U, W and X lists that M2L does not need.
    bool prep_done = false, m2l_launched = false;
    auto on_vlists = [&](const TreeView& t, const TreeMeta&) {
        run_m2l_prep(c->pool, s, t, c->prm.p, c->prm.precision, c->shapes);
        prep_done = true;
        if (lists_beside) {
            AFMM_CHECK(hipStreamWaitEvent(s, c->up_done, 0));  // the multipoles are done
            run_m2l(c->pool, s, t, c->prm.p, tab, c->prm.precision, c->shapes, c->ev[7], nullptr);
            AFMM_CHECK(hipEventRecord(c->m2l_done, s));
            m2l_launched = true;
        }
    };
    build_forest(c->pool, s, pos, mass, nsys, sys_off, eps, c->prm.leaf_cap, tv, meta, roots,
                 split ? TreeHook(on_core) : TreeHook(), split ? TreeHook(on_vlists) : TreeHook(),
                 lists_beside ? c->side : nullptr, c->list_fork, &c->key_levels_hint, &c->tree_depth_hint);
    for (const RootInfo& r : roots)
        if (r.nonfinite) return fail(c, AFMM_ERR_NONFINITE, "position or mass is NaN or Inf");
    if (lists_beside) {
        // the tree phase ends with the lists on the second stream; P2P follows
        // them there (in the automatic schedule after the bf16 M2L, whose short
        // M2L it slowed on the 1M merger; AFMM_OVERLAP_BESIDE_M2L keeps it
        // beside, for near fields as large as the far field)
        AFMM_CHECK(hipEventRecord(c->ev[1], c->side));
        AFMM_CHECK(hipEventRecord(c->lists_done, c->side));
        // P2L and M2P (needed first, by L2L and L2P on s), then P2P
        run_far_passes(c->pool, c->side, tv, c->prm.p, phi != nullptr, c->x_done, c->w_done);
        if (bf16_m2l && c->overlap_mode != AFMM_OVERLAP_BESIDE_M2L)
            AFMM_CHECK(hipStreamWaitEvent(c->side, c->m2l_done, 0));
        run_p2p_near(c->pool, c->side, tv, c->prm.precision, c->shapes, phi != nullptr, c->ev[8], c->ev[9]);
        AFMM_CHECK(hipEventRecord(c->side_join, c->side));
    } else {
        AFMM_CHECK(hipEventRecord(c->ev[1], s));
    }
    if (c->use_graph) {
        // the upward pass and the M2L preparation already ran outside the graph
        // (beside and within the list build); the stream waits for them before
        // the capture
        if (up_launched) AFMM_CHECK(hipStreamWaitEvent(s, c->up_done, 0));
        SideStream gside{};
        gside.upward_launched = up_launched;
        gside.prep_done = prep_done;
        // P2P on a second stream inside the graph too (a branch of the graph
        // beside the far-field passes); the capture takes the context's own
        // side stream, so no other context's work can join it
        if (split) {
            if (!c->graph_side) AFMM_CHECK(hipStreamCreateWithFlags(&c->graph_side, hipStreamNonBlocking));
            gside.s = c->graph_side;
            gside.fork = c->side_fork;
            gside.join = c->side_join;
        }
        // size every buffer first, so that the capture allocates nothing
        run_fmm(c->pool, s, tv, meta, c->prm.p, tab, c->prm.precision, c->shapes, acc, phi, nullptr, true, gside);
        hipGraph_t g = nullptr;
        AFMM_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            run_fmm(c->pool, s, tv, meta, c->prm.p, tab, c->prm.precision, c->shapes, acc, phi, nullptr, false, gside);
        } catch (...) {
            if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
            throw;
        }
        AFMM_CHECK(hipStreamEndCapture(s, &g));
        bool updated = false;
        if (c->gexec) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            updated = hipGraphExecUpdate(c->gexec, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
            if (!updated) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(c->gexec);
                c->gexec = nullptr;
            }
        }
        if (!c->gexec) {
            const hipError_t e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(g);
                c->gexec = nullptr;
                throw HipError(e, "hipGraphInstantiate");
            }
        }
        AFMM_CHECK(hipGraphDestroy(g));
        AFMM_CHECK(hipGraphLaunch(c->gexec, s));
        AFMM_CHECK(hipEventRecord(c->ev[6], s));
        (updated ? c->graph_updates : c->graph_instantiations) += 1;
    } else {
        SideStream side{};
        if (split) side = SideStream{c->side, c->side_fork, c->side_join, up_launched, c->up_done, prep_done};
        side.m2l_launched = m2l_launched;
        side.p2p_launched = lists_beside;
        side.lists_done = lists_beside ? c->lists_done : nullptr;
        side.far_beside = lists_beside;
        side.x_done = c->x_done;
        side.w_done = c->w_done;
        run_fmm(c->pool, s, tv, meta, c->prm.p, tab, c->prm.precision, c->shapes, acc, phi, c->ev + 2, false, side);
    }
    AFMM_CHECK(hipStreamSynchronize(s));
    c->tv = tv;
    c->meta = meta;
    c->root = roots[0];
    c->nsys = nsys;
    c->have_tree = true;
    afmm_stats& st = c->stats;
    st = afmm_stats{};
    st.n_bodies = n;
    st.n_boxes = tv.nb;
    st.n_leaves = tv.nleaves;
    st.n_u = tv.n_u;
    st.n_w = tv.n_w;
    st.n_x = tv.n_x;
    st.max_level = tv.max_level;
    unsigned long long tot[4];
    AFMM_CHECK(hipMemcpy(tot, tv.list_totals, sizeof(tot), hipMemcpyDeviceToHost));
    st.n_colleague = (int64_t)tot[0];
    st.n_v = (int64_t)tot[1];
    st.max_leaf_occ = (int32_t)tot[2];
    st.p2p_pairs = (int64_t)tot[3] - n;
    st.ms_tree = elapsed(c->ev[0], c->ev[1]);
    st.ms_total = elapsed(c->ev[0], c->ev[6]);
    if (!c->use_graph) {  // the graph has no per-phase events
        st.ms_upward = elapsed(c->ev[2], c->ev[3]);
        st.ms_lists = elapsed(c->ev[3], c->ev[7]);  // M2L ordering / source table
        st.ms_m2l = elapsed(c->ev[7], c->ev[4]);
        st.ms_downward = elapsed(c->ev[4], c->ev[5]);
        st.ms_leaf = elapsed(c->ev[5], c->ev[6]);
        st.ms_p2p = elapsed(c->ev[8], c->ev[9]);
    }
    st.graph_instantiations = c->graph_instantiations;
    st.graph_updates = c->graph_updates;
    return AFMM_OK;
}

namespace {

// Validates a batch description; fills 32-bit offsets and per-system G, eps.
int batch_args(afmm_ctx* c, int32_t nsys, const int64_t* sys_off, const double* G, const double* eps,
               std::vector<int>& off, std::vector<double>& g, std::vector<double>& e) {
    if (nsys < 1 || !sys_off || sys_off[0] != 0) return fail(c, AFMM_ERR_ARG, "bad system offsets");
    if (sys_off[nsys] >= (1LL << 30)) return fail(c, AFMM_ERR_TOO_LARGE, "n beyond 2^30");
    off.resize(nsys + 1);
    g.resize(nsys);
    e.resize(nsys);
    for (int q = 0; q <= nsys; ++q) {
        if (q > 0 && sys_off[q] <= sys_off[q - 1]) return fail(c, AFMM_ERR_ARG, "every system needs a body");
        off[q] = (int)sys_off[q];
    }
    for (int q = 0; q < nsys; ++q) {
        g[q] = G ? G[q] : c->prm.G;
        e[q] = eps ? eps[q] : c->prm.eps;
        if (!std::isfinite(g[q]) || !std::isfinite(e[q]) || e[q] < 0)
            return fail(c, AFMM_ERR_ARG, "bad per-system G or eps");
    }
    return AFMM_OK;
}

}  // namespace

extern "C" {

const char* afmm_strerror(int code) {
    switch (code) {
        case AFMM_OK: return "ok";
        case AFMM_ERR_ARG: return "invalid argument";
        case AFMM_ERR_ORDER: return "expansion order outside [1, 16]";
        case AFMM_ERR_LEAF_CAP: return "leaf capacity must be >= 1";
        case AFMM_ERR_NONFINITE: return "position or mass is NaN or Inf";
        case AFMM_ERR_DEVICE: return "HIP runtime error";
        case AFMM_ERR_NO_MEMORY: return "device memory exhausted";
        case AFMM_ERR_STATE: return "no evaluation has run on this context";
        case AFMM_ERR_TOO_LARGE: return "too many bodies for 32-bit indexing";
        case AFMM_ERR_PRECISION: return "precision mode not supported";
    }
    return "unknown error";
}

const char* afmm_last_error(const afmm_ctx* c) { return c ? c->err.c_str() : ""; }

int afmm_create(const afmm_params* prm, afmm_ctx** out) {
    if (!out) return AFMM_ERR_ARG;
    *out = nullptr;
    const int rc = check_params(prm);
    if (rc) return rc;
    afmm_ctx* c = new (std::nothrow) afmm_ctx();
    if (!c) return AFMM_ERR_NO_MEMORY;
    c->prm = *prm;
    const int r = guarded(c, [&]() -> int {
        AFMM_CHECK(hipSetDevice(prm->device));
        direct_scratch_setup(prm->device);  // once per device: the static scratch's address
        c->shapes.m2l_waves = default_m2l_waves(prm->p, prm->precision);
        AFMM_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (auto& e : c->ev) AFMM_CHECK(hipEventCreate(&e));
        const Tables& T = tables_for_order(prm->p);
        const std::vector<double> m2m = mfma_a_layout(T.m2m, T.S, 8);
        const std::vector<double> l2l = mfma_a_layout(T.l2l, T.S, 8);
        // M2L operators in slot order (the order the M2L kernels step through them)
        std::vector<double> m2l_slots(T.m2l.size());
        const size_t SS = (size_t)T.S * T.S;
        for (int k = 0; k < kNumOffsets; ++k)
            std::copy(T.m2l.begin() + kM2LSlotsHost.offset_of[k] * SS,
                      T.m2l.begin() + (kM2LSlotsHost.offset_of[k] + 1) * SS, m2l_slots.begin() + k * SS);
        const std::vector<double> m2l = m2l_a_layout(m2l_slots, T.S, kNumOffsets);
        c->d_m2m_sw = c->pool.get<double>("T_m2m_sw", m2m.size());
        c->d_l2l_sw = c->pool.get<double>("T_l2l_sw", l2l.size());
        c->d_m2l_sw = c->pool.get<double>("T_m2l_sw", m2l.size());
        AFMM_CHECK(hipMemcpy(c->d_m2m_sw, m2m.data(), sizeof(double) * m2m.size(), hipMemcpyHostToDevice));
        AFMM_CHECK(hipMemcpy(c->d_l2l_sw, l2l.data(), sizeof(double) * l2l.size(), hipMemcpyHostToDevice));
        AFMM_CHECK(hipMemcpy(c->d_m2l_sw, m2l.data(), sizeof(double) * m2l.size(), hipMemcpyHostToDevice));
        if (m2l_form(prm->precision) == kPrecMixed) {
            std::vector<float2> hl(m2l.size());
            for (size_t i = 0; i < m2l.size(); ++i) {
                const float hi = (float)m2l[i];
                hl[i] = float2{hi, (float)(m2l[i] - (double)hi)};
            }
            c->d_m2l_hl = c->pool.get<float2>("T_m2l_hl", hl.size());
            AFMM_CHECK(hipMemcpy(c->d_m2l_hl, hl.data(), sizeof(float2) * hl.size(), hipMemcpyHostToDevice));
        }
        if (m2l_form(prm->precision) == kPrecMixedBf16) {
            const std::vector<uint16_t> bf = mfma_bf16_layout(m2l_slots, T.S, kNumOffsets);
            c->d_m2l_bf = c->pool.get<uint4>("T_m2l_bf", bf.size() / 8);
            AFMM_CHECK(hipMemcpy(c->d_m2l_bf, bf.data(), sizeof(uint16_t) * bf.size(), hipMemcpyHostToDevice));
        }
        return AFMM_OK;
    });
    if (r) {
        afmm_destroy(c);
        return r;
    }
    *out = c;
    return AFMM_OK;
}

int afmm_set_graph(afmm_ctx* c, int enable) {
    if (!c) return AFMM_ERR_ARG;
    c->use_graph = enable != 0;
    return AFMM_OK;
}

int afmm_set_overlap(afmm_ctx* c, int mode) {
    if (!c) return AFMM_ERR_ARG;
    if (mode < AFMM_OVERLAP_OFF || mode > AFMM_OVERLAP_IN_ORDER)
        return fail(c, AFMM_ERR_ARG, "overlap mode must be 0 (off), 1 (auto), 2 (beside M2L) or 3 (in order)");
    c->overlap_mode = mode;
    c->split_p2p = mode != AFMM_OVERLAP_OFF;
    return AFMM_OK;
}

void afmm_destroy(afmm_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipEventSynchronize(c->side_join);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->pool.release();
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {c->side_fork, c->side_join, c->up_fork, c->up_done, c->list_fork, c->lists_done, c->m2l_done,
                         c->x_done, c->w_done})
        if (e) (void)hipEventDestroy(e);
    if (c->graph_side) (void)hipStreamDestroy(c->graph_side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int afmm_synchronize(afmm_ctx* c) {
    if (!c) return AFMM_ERR_ARG;
    return guarded(c, [&]() -> int {
        AFMM_CHECK(hipStreamSynchronize(c->stream));
        return AFMM_OK;
    });
}

int afmm_eval_device_potential(afmm_ctx* c, int64_t n, const double* pos, const double* mass, double* acc,
                               double* phi) {
    if (!c || n < 0 || (n > 0 && (!pos || !mass || !acc))) return AFMM_ERR_ARG;
    if (n >= (1LL << 30)) return fail(c, AFMM_ERR_TOO_LARGE, "n beyond 2^30");
    return guarded(c, [&]() -> int {
        AFMM_CHECK(hipSetDevice(c->prm.device));
        const int off[2] = {0, (int)n};
        return eval_forest(c, 1, off, &c->prm.G, &c->prm.eps, pos, mass, acc, phi);
    });
}

int afmm_eval_device(afmm_ctx* c, int64_t n, const double* pos, const double* mass, double* acc) {
    return afmm_eval_device_potential(c, n, pos, mass, acc, nullptr);
}

int afmm_eval_potential(afmm_ctx* c, int64_t n, const double* pos, const double* mass, double* acc, double* phi) {
    if (!c || n < 0 || (n > 0 && (!pos || !mass || !acc))) return AFMM_ERR_ARG;
    if (n >= (1LL << 30)) return fail(c, AFMM_ERR_TOO_LARGE, "n beyond 2^30");
    if (n == 0) {
        c->have_tree = false;
        return AFMM_OK;
    }
    return guarded(c, [&]() -> int {
        AFMM_CHECK(hipSetDevice(c->prm.device));
        hipStream_t s = c->stream;
        double* dpos = c->pool.get<double>("in_pos", 3 * (size_t)n);
        double* dmass = c->pool.get<double>("in_mass", n);
        double* dacc = c->pool.get<double>("out_acc", 3 * (size_t)n);
        double* dphi = phi ? c->pool.get<double>("out_phi", n) : nullptr;
        c->stage.upload(dpos, pos, sizeof(double) * 3 * n, s);
        c->stage.upload(dmass, mass, sizeof(double) * n, s);
        const int off[2] = {0, (int)n};
        const int rc = eval_forest(c, 1, off, &c->prm.G, &c->prm.eps, dpos, dmass, dacc, dphi);
        if (rc) return rc;
        c->stage.download(acc, dacc, sizeof(double) * 3 * n, s);
        if (phi) c->stage.download(phi, dphi, sizeof(double) * n, s);
        return AFMM_OK;
    });
}

int afmm_eval(afmm_ctx* c, int64_t n, const double* pos, const double* mass, double* acc) {
    return afmm_eval_potential(c, n, pos, mass, acc, nullptr);
}

int afmm_eval_batch_device(afmm_ctx* c, int32_t nsys, const int64_t* sys_off, const double* G, const double* eps,
                           const double* pos, const double* mass, double* acc, double* phi) {
    if (!c || !pos || !mass || !acc) return AFMM_ERR_ARG;
    std::vector<int> off;
    std::vector<double> g, e;
    if (int rc = batch_args(c, nsys, sys_off, G, eps, off, g, e)) return rc;
    return guarded(c, [&]() -> int {
        AFMM_CHECK(hipSetDevice(c->prm.device));
        return eval_forest(c, nsys, off.data(), g.data(), e.data(), pos, mass, acc, phi);
    });
}

int afmm_eval_batch(afmm_ctx* c, int32_t nsys, const int64_t* sys_off, const double* G, const double* eps,
                    const double* pos, const double* mass, double* acc, double* phi) {
    if (!c || !pos || !mass || !acc) return AFMM_ERR_ARG;
    std::vector<int> off;
    std::vector<double> g, e;
    if (int rc = batch_args(c, nsys, sys_off, G, eps, off, g, e)) return rc;
    const int64_t n = off[nsys];
    return guarded(c, [&]() -> int {
        AFMM_CHECK(hipSetDevice(c->prm.device));
        hipStream_t s = c->stream;
        double* dpos = c->pool.get<double>("in_pos", 3 * (size_t)n);
        double* dmass = c->pool.get<double>("in_mass", n);
        double* dacc = c->pool.get<double>("out_acc", 3 * (size_t)n);
        double* dphi = phi ? c->pool.get<double>("out_phi", n) : nullptr;
        c->stage.upload(dpos, pos, sizeof(double) * 3 * n, s);
        c->stage.upload(dmass, mass, sizeof(double) * n, s);
        const int rc = eval_forest(c, nsys, off.data(), g.data(), e.data(), dpos, dmass, dacc, dphi);
        if (rc) return rc;
        c->stage.download(acc, dacc, sizeof(double) * 3 * n, s);
        if (phi) c->stage.download(phi, dphi, sizeof(double) * n, s);
        return AFMM_OK;
    });
}

int afmm_direct(afmm_ctx* c, int64_t n, const double* pos, const double* mass, int64_t ntgt, const int64_t* tgt,
                double* acc, double* phi) {
    if (!c || n < 0 || ntgt < 0) return AFMM_ERR_ARG;
    if (ntgt == 0) return AFMM_OK;
    if (!pos || !mass || !tgt || !acc) return AFMM_ERR_ARG;
    if (n >= (1LL << 30) || ntgt >= (1LL << 30)) return fail(c, AFMM_ERR_TOO_LARGE, "n beyond 2^30");
    std::vector<int> t32(ntgt);
    for (int64_t k = 0; k < ntgt; ++k) {
        if (tgt[k] < 0 || tgt[k] >= n) return fail(c, AFMM_ERR_ARG, "target index outside [0, n)");
        t32[k] = (int)tgt[k];
    }
    return guarded(c, [&]() -> int {
        AFMM_CHECK(hipSetDevice(c->prm.device));
        hipStream_t s = c->stream;
        double* dpos = c->pool.get<double>("dir_pos", 3 * (size_t)n);
        double* dmass = c->pool.get<double>("dir_mass", n);
        int* dtgt = c->pool.get<int>("dir_tgt", ntgt);
        double* dacc = c->pool.get<double>("dir_acc", 3 * (size_t)ntgt);
        double* dphi = c->pool.get<double>("dir_phi", ntgt);
        AFMM_CHECK(hipMemcpyAsync(dpos, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
        AFMM_CHECK(hipMemcpyAsync(dmass, mass, sizeof(double) * n, hipMemcpyHostToDevice, s));
        AFMM_CHECK(hipMemcpyAsync(dtgt, t32.data(), sizeof(int) * ntgt, hipMemcpyHostToDevice, s));
        direct_targets(c->pool, s, dpos, dmass, (int)n, dtgt, (int)ntgt, c->prm.G, c->prm.eps, dacc, dphi);
        AFMM_CHECK(hipMemcpyAsync(acc, dacc, sizeof(double) * 3 * ntgt, hipMemcpyDeviceToHost, s));
        if (phi) AFMM_CHECK(hipMemcpyAsync(phi, dphi, sizeof(double) * ntgt, hipMemcpyDeviceToHost, s));
        AFMM_CHECK(hipStreamSynchronize(s));
        return AFMM_OK;
    });
}

int afmm_direct_device_noalloc(afmm_ctx* c, int64_t n, const double* d_pos, const double* d_mass, int64_t ntgt,
                               const int64_t* d_tgt, double* d_acc, double* d_phi, void* stream) {
    if (!c || n < 0 || ntgt < 0) return AFMM_ERR_ARG;
    if (ntgt == 0) return AFMM_OK;
    if (!d_pos || !d_mass || !d_tgt || !d_acc) return AFMM_ERR_ARG;
    if (n >= (1LL << 30)) return fail(c, AFMM_ERR_TOO_LARGE, "n beyond 2^30");
    if (ntgt > AFMM_DIRECT_NOALLOC_MAX_TARGETS)
        return fail(c, AFMM_ERR_TOO_LARGE, "ntgt beyond AFMM_DIRECT_NOALLOC_MAX_TARGETS (static scratch)");
    return guarded(c, [&]() -> int {
        AFMM_CHECK(hipSetDevice(c->prm.device));
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
        direct_targets_noalloc(c->prm.device, s, d_pos, d_mass, (int)n, d_tgt, (int)ntgt, c->prm.G, c->prm.eps, d_acc,
                               d_phi);
        return AFMM_OK;
    });
}

int afmm_set_shapes(afmm_ctx* c, const afmm_shapes* sh) {
    if (!c) return AFMM_ERR_ARG;
    Shapes v{};
    v.m2l_waves = default_m2l_waves(c->prm.p, c->prm.precision);
    if (sh) {
        if (sh->m2l_waves) {
            if (sh->m2l_waves != 8 && sh->m2l_waves != 16) return fail(c, AFMM_ERR_ARG, "m2l_waves must be 8 or 16");
            v.m2l_waves = sh->m2l_waves;
        }
        if (sh->p2p_pair_bodies < 0) return fail(c, AFMM_ERR_ARG, "p2p_pair_bodies must be >= 0");
        if (sh->p2p_pair_bodies) v.p2p_pair_bodies = sh->p2p_pair_bodies;
        const int g = sh->p2p_min_group;
        if (g < 0 || g > 64 || (g & (g - 1))) return fail(c, AFMM_ERR_ARG, "p2p_min_group must be 0 or a power of two <= 64");
        if (g) v.p2p_min_group = g;
        if (sh->p2p_schedule < 0 || sh->p2p_schedule > 2) return fail(c, AFMM_ERR_ARG, "p2p_schedule must be 0, 1 or 2");
        v.p2p_schedule = sh->p2p_schedule == 1 ? 1 : 0;
    }
    c->shapes = v;
    return AFMM_OK;
}

int afmm_get_shapes(const afmm_ctx* c, afmm_shapes* out) {
    if (!c || !out) return AFMM_ERR_ARG;
    out->m2l_waves = c->shapes.m2l_waves;
    out->p2p_pair_bodies = c->shapes.p2p_pair_bodies;
    out->p2p_min_group = c->shapes.p2p_min_group;
    out->p2p_schedule = c->shapes.p2p_schedule ? 1 : 2;
    return AFMM_OK;
}

int afmm_get_stats(const afmm_ctx* c, afmm_stats* out) {
    if (!c || !out) return AFMM_ERR_ARG;
    *out = c->stats;
    return AFMM_OK;
}

// Reference numbering: box ids by allocation order (refid), lists translated
// element-wise; their orders already match the reference.
int afmm_tree_export(afmm_ctx* c, afmm_tree* out) {
    if (!c || !out) return AFMM_ERR_ARG;
    if (!c->have_tree) return fail(c, AFMM_ERR_STATE, "no tree: run afmm_eval first");
    if (c->nsys != 1) return fail(c, AFMM_ERR_STATE, "tree export covers single-system evaluations");
    return guarded(c, [&]() -> int {
        const TreeView& t = c->tv;
        const int nb = t.nb, n = t.n, nl = t.nleaves;
        auto d2h = [](std::vector<int>& h, const int* d, size_t cnt) {
            h.resize(cnt);
            if (cnt) AFMM_CHECK(hipMemcpy(h.data(), d, sizeof(int) * cnt, hipMemcpyDeviceToHost));
        };
        auto d2hd = [](std::vector<double>& h, const double* d, size_t cnt) {
            h.resize(cnt);
            if (cnt) AFMM_CHECK(hipMemcpy(h.data(), d, sizeof(double) * cnt, hipMemcpyDeviceToHost));
        };
        std::vector<int> refid, parent, level, cfirst, nchild, octant, sfirst, scount, lids, coll, ncoll, vl, nv,
            uoff, uidx, woff, widx, xoff, xidx, orig;
        std::vector<double> cfp, half, x, y, z, m;
        d2h(refid, t.refid, nb);
        d2h(parent, t.parent, nb);
        d2h(level, t.level, nb);
        d2h(cfirst, t.child_first, nb);
        d2h(nchild, t.nchild, nb);
        d2h(octant, t.octant, nb);
        d2h(sfirst, t.seg_first, nb);
        d2h(scount, t.seg_count, nb);
        d2h(lids, t.leaf_ids, nl);
        d2h(coll, t.coll, (size_t)kCollSlots * nb);
        d2h(ncoll, t.ncoll, nb);
        d2h(vl, t.vl, (size_t)kVSlots * nb);
        d2h(nv, t.nv, nb);
        d2h(uoff, t.u_off, nl + 1);
        d2h(uidx, t.u_idx, t.n_u);
        d2h(woff, t.w_off, nl + 1);
        d2h(widx, t.w_idx, t.n_w);
        d2h(xoff, t.x_off, nb + 1);
        d2h(xidx, t.x_idx, t.n_x);
        d2h(orig, t.orig, n);
        d2hd(cfp, t.cfp, 3 * (size_t)nb);
        d2hd(half, t.half, nb);
        d2hd(x, t.x, n);
        d2hd(y, t.y, n);
        d2hd(z, t.z, n);
        d2hd(m, t.m, n);
        std::vector<int> bfs_of(nb), leaf_rank(nb, -1);
        for (int b = 0; b < nb; ++b) bfs_of[refid[b]] = b;
        for (int li = 0; li < nl; ++li) leaf_rank[lids[li]] = li;
        for (int r = 0; r < nb; ++r) {
            const int b = bfs_of[r];
            if (out->parent) out->parent[r] = parent[b] < 0 ? -1 : refid[parent[b]];
            if (out->child)
                for (int o = 0; o < 8; ++o) out->child[8 * r + o] = -1;
            if (out->level) out->level[r] = level[b];
            if (out->center)
                for (int a = 0; a < 3; ++a) out->center[3 * r + a] = cfp[3 * (size_t)b + a];
            if (out->half) out->half[r] = half[b];
            if (out->leaf) out->leaf[r] = nchild[b] == 0;
            if (out->pfirst) out->pfirst[r] = sfirst[b];
            if (out->pcount) out->pcount[r] = scount[b];
        }
        if (out->child)
            for (int b = 1; b < nb; ++b) out->child[8 * refid[parent[b]] + octant[b]] = refid[b];
        for (int k = 0; k < n; ++k) {
            if (out->px) out->px[k] = x[k];
            if (out->py) out->py[k] = y[k];
            if (out->pz) out->pz[k] = z[k];
            if (out->mass) out->mass[k] = m[k];
            if (out->orig_index) out->orig_index[k] = orig[k];
        }
        if (out->leaves) {
            std::vector<int> lv;
            for (int li = 0; li < nl; ++li) lv.push_back(refid[lids[li]]);
            std::sort(lv.begin(), lv.end());
            std::copy(lv.begin(), lv.end(), out->leaves);
        }
        // CSR in reference box order
        auto emit = [&](int32_t* off, int32_t* idx, auto count_of, auto item_of) {
            int64_t k = 0;
            if (off) off[0] = 0;
            for (int r = 0; r < nb; ++r) {
                const int b = bfs_of[r];
                const int cnt = count_of(b);
                for (int e = 0; e < cnt; ++e) {
                    if (idx) idx[k] = refid[item_of(b, e)];
                    ++k;
                }
                if (off) off[r + 1] = (int32_t)k;
            }
        };
        emit(out->colleague_off, out->colleague_idx, [&](int b) { return ncoll[b]; },
             [&](int b, int e) { return coll[(size_t)kCollSlots * b + e]; });
        emit(out->v_off, out->v_idx, [&](int b) { return nv[b]; },
             [&](int b, int e) { return vl[(size_t)kVSlots * b + e]; });
        emit(out->u_off, out->u_idx,
             [&](int b) { return leaf_rank[b] < 0 ? 0 : uoff[leaf_rank[b] + 1] - uoff[leaf_rank[b]]; },
             [&](int b, int e) { return uidx[uoff[leaf_rank[b]] + e]; });
        emit(out->w_off, out->w_idx,
             [&](int b) { return leaf_rank[b] < 0 ? 0 : woff[leaf_rank[b] + 1] - woff[leaf_rank[b]]; },
             [&](int b, int e) { return widx[woff[leaf_rank[b]] + e]; });
        emit(out->x_off, out->x_idx, [&](int b) { return xoff[b + 1] - xoff[b]; },
             [&](int b, int e) { return xidx[xoff[b] + e]; });
        out->tol = t.tol;
        return AFMM_OK;
    });
}

// ---- resident leapfrog --------------------------------------------------------
struct afmm_state {
    afmm_ctx* ctx = nullptr;
    int nsys = 0;
    std::vector<int> off;
    std::vector<double> G, eps;
    Pool mem;  // the state's own arrays (released with the state)
    double *pos = nullptr, *vel = nullptr, *mass = nullptr, *acc = nullptr, *phi = nullptr;
    std::vector<JobRef> jobs;
    JobRef* d_jobs = nullptr;
    int *d_boff = nullptr, *d_n = nullptr;
    int64_t steps = 0;
};

namespace {
// One batched evaluation of the state's systems, with drift/kick when `step`.
int state_eval(afmm_state* st, bool step, double dt, bool with_phi) {
    afmm_ctx* c = st->ctx;
    hipStream_t s = c->stream;
    const int n = st->off[st->nsys];
    for (JobRef& j : st->jobs) {
        j.dt = dt;
        j.drift = step ? 1 : 0;
        j.phi = with_phi ? j.phi : nullptr;
    }
    if (with_phi)
        for (int q = 0; q < st->nsys; ++q) st->jobs[q].phi = st->phi + st->off[q];
    AFMM_CHECK(hipMemcpyAsync(st->d_jobs, st->jobs.data(), sizeof(JobRef) * st->nsys, hipMemcpyHostToDevice, s));
    double* bpos = c->pool.get<double>("lf_pos", 3 * (size_t)n);
    double* bmass = c->pool.get<double>("lf_mass", n);
    double* bacc = c->pool.get<double>("lf_acc", 3 * (size_t)n);
    double* bphi = with_phi ? c->pool.get<double>("lf_phi", n) : nullptr;
    leapfrog_gather(s, n, st->nsys, st->d_boff, st->d_jobs, bpos, bmass);
    const int rc = eval_forest(c, st->nsys, st->off.data(), st->G.data(), st->eps.data(), bpos, bmass, bacc, bphi);
    if (rc) return rc;
    leapfrog_kick(s, n, st->nsys, st->d_boff, st->d_jobs, bacc, bphi);
    for (JobRef& j : st->jobs) j.init = 1;
    AFMM_CHECK(hipMemcpyAsync(st->d_jobs, st->jobs.data(), sizeof(JobRef) * st->nsys, hipMemcpyHostToDevice, s));
    AFMM_CHECK(hipStreamSynchronize(s));
    return AFMM_OK;
}
}  // namespace

int afmm_state_create(afmm_ctx* c, int32_t nsys, const int64_t* sys_off, const double* G, const double* eps,
                      const double* pos, const double* vel, const double* mass, afmm_state** out) {
    if (!c || !out || !pos || !vel || !mass) return AFMM_ERR_ARG;
    *out = nullptr;
    std::vector<int> off;
    std::vector<double> g, e;
    if (int rc = batch_args(c, nsys, sys_off, G, eps, off, g, e)) return rc;
    afmm_state* st = new (std::nothrow) afmm_state();
    if (!st) return AFMM_ERR_NO_MEMORY;
    st->ctx = c;
    st->nsys = nsys;
    st->off = off;
    st->G = g;
    st->eps = e;
    const int n = off[nsys];
    const int rc = guarded(c, [&]() -> int {
        AFMM_CHECK(hipSetDevice(c->prm.device));
        hipStream_t s = c->stream;
        st->pos = st->mem.get<double>("pos", 3 * (size_t)n);
        st->vel = st->mem.get<double>("vel", 3 * (size_t)n);
        st->acc = st->mem.get<double>("acc", 3 * (size_t)n);
        st->mass = st->mem.get<double>("mass", n);
        st->phi = st->mem.get<double>("phi", n);
        st->d_jobs = st->mem.get<JobRef>("jobs", nsys);
        st->d_boff = st->mem.get<int>("boff", nsys + 1);
        st->d_n = st->mem.get<int>("n", nsys);
        AFMM_CHECK(hipMemcpyAsync(st->pos, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
        AFMM_CHECK(hipMemcpyAsync(st->vel, vel, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
        AFMM_CHECK(hipMemcpyAsync(st->mass, mass, sizeof(double) * n, hipMemcpyHostToDevice, s));
        std::vector<int> cnt(nsys);
        for (int q = 0; q < nsys; ++q) cnt[q] = off[q + 1] - off[q];
        AFMM_CHECK(hipMemcpyAsync(st->d_boff, off.data(), sizeof(int) * (nsys + 1), hipMemcpyHostToDevice, s));
        AFMM_CHECK(hipMemcpyAsync(st->d_n, cnt.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, s));
        AFMM_CHECK(hipStreamSynchronize(s));
        st->jobs.resize(nsys);
        for (int q = 0; q < nsys; ++q) {
            const size_t o = (size_t)off[q];
            st->jobs[q] = JobRef{st->pos + 3 * o, st->vel + 3 * o, st->acc + 3 * o, nullptr, st->mass + o, 0.0, 0, 0};
        }
        return state_eval(st, false, 0.0, false);  // VelocityVerlet3D::initialize
    });
    if (rc) {
        delete st;
        return rc;
    }
    *out = st;
    return AFMM_OK;
}

int afmm_state_step(afmm_state* st, double dt, int32_t nsteps) {
    if (!st || nsteps < 0 || !std::isfinite(dt)) return AFMM_ERR_ARG;
    afmm_ctx* c = st->ctx;
    return guarded(c, [&]() -> int {
        AFMM_CHECK(hipSetDevice(c->prm.device));
        for (int k = 0; k < nsteps; ++k) {
            if (int rc = state_eval(st, true, dt, false)) return rc;
            ++st->steps;
        }
        return AFMM_OK;
    });
}

int afmm_state_download(afmm_state* st, double* pos, double* vel, double* acc) {
    if (!st) return AFMM_ERR_ARG;
    afmm_ctx* c = st->ctx;
    const size_t n = (size_t)st->off[st->nsys];
    return guarded(c, [&]() -> int {
        AFMM_CHECK(hipSetDevice(c->prm.device));
        hipStream_t s = c->stream;
        if (pos) AFMM_CHECK(hipMemcpyAsync(pos, st->pos, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
        if (vel) AFMM_CHECK(hipMemcpyAsync(vel, st->vel, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
        if (acc) AFMM_CHECK(hipMemcpyAsync(acc, st->acc, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
        AFMM_CHECK(hipStreamSynchronize(s));
        return AFMM_OK;
    });
}

int afmm_state_diagnostics(afmm_state* st, double* out) {
    if (!st || !out) return AFMM_ERR_ARG;
    afmm_ctx* c = st->ctx;
    return guarded(c, [&]() -> int {
        AFMM_CHECK(hipSetDevice(c->prm.device));
        if (int rc = state_eval(st, false, 0.0, true)) return rc;
        hipStream_t s = c->stream;
        double* d_out = c->pool.get<double>("diag_out", 9 * (size_t)st->nsys);
        diagnostics(s, st->nsys, st->d_jobs, st->d_n, d_out);
        AFMM_CHECK(hipMemcpyAsync(out, d_out, sizeof(double) * 9 * st->nsys, hipMemcpyDeviceToHost, s));
        AFMM_CHECK(hipStreamSynchronize(s));
        return AFMM_OK;
    });
}

void afmm_state_destroy(afmm_state* st) {
    if (!st) return;
    if (st->ctx && st->ctx->stream) (void)hipStreamSynchronize(st->ctx->stream);
    delete st;
}

}  // extern "C"
