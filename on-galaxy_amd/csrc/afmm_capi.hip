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
#include <new>
#include <string>
#include <vector>

#include "../../include/afmm.h"
#include "afmm_internal.h"
#include "afmm_math.h"
#include "afmm_tables.h"

using namespace afmm;

struct afmm_ctx {
    afmm_params prm{};
    hipStream_t stream = nullptr;
    Pool pool;
    std::string err;
    // tables
    double *d_m2m_sw = nullptr, *d_l2l_sw = nullptr, *d_m2l_sw = nullptr;
    // last evaluation
    bool have_tree = false;
    TreeView tv{};
    TreeMeta meta;
    RootInfo root{};
    afmm_stats stats{};
    hipEvent_t ev[8] = {};
};

namespace {

int fail(afmm_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

template <class F>
int guarded(afmm_ctx* c, F&& f) {
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

int check_params(const afmm_params* p) {
    if (!p) return AFMM_ERR_ARG;
    if (!(std::isfinite(p->G)) || !(std::isfinite(p->eps)) || p->eps < 0) return AFMM_ERR_ARG;
    if (p->p < 1 || p->p > kMaxOrder) return AFMM_ERR_ORDER;
    if (p->leaf_cap < 1) return AFMM_ERR_LEAF_CAP;
    if (p->precision != AFMM_PRECISION_FP64) return AFMM_ERR_PRECISION;
    return AFMM_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

// The evaluation on device arrays (pos AoS n x 3, mass n) into acc (AoS, input order).
int eval_device(afmm_ctx* c, int64_t n, const double* pos, const double* mass, double* acc) {
    c->have_tree = false;
    if (n == 0) return AFMM_OK;
    hipStream_t s = c->stream;
    AFMM_CHECK(hipEventRecord(c->ev[0], s));
    TreeView tv{};
    TreeMeta meta;
    RootInfo root{};
    build_tree(c->pool, s, pos, mass, (int)n, c->prm.leaf_cap, c->prm.eps, tv, meta, root);
    if (root.nonfinite) return fail(c, AFMM_ERR_NONFINITE, "position or mass is NaN or Inf");
    AFMM_CHECK(hipEventRecord(c->ev[1], s));
    FmmTables tab{c->d_m2m_sw, c->d_l2l_sw, c->d_m2l_sw};
    run_fmm(c->pool, s, tv, meta, c->prm.p, tab, c->prm.G, c->prm.eps, acc, c->ev + 2);
    AFMM_CHECK(hipStreamSynchronize(s));
    c->tv = tv;
    c->meta = meta;
    c->root = root;
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
    // small host reductions for the stats
    std::vector<int> nc(tv.nb), nv(tv.nb), cnt(tv.nleaves), lid(tv.nleaves), scount(tv.nb), uoff(tv.nleaves + 1),
        uidx(tv.n_u);
    AFMM_CHECK(hipMemcpy(nc.data(), tv.ncoll, sizeof(int) * tv.nb, hipMemcpyDeviceToHost));
    AFMM_CHECK(hipMemcpy(nv.data(), tv.nv, sizeof(int) * tv.nb, hipMemcpyDeviceToHost));
    AFMM_CHECK(hipMemcpy(lid.data(), tv.leaf_ids, sizeof(int) * tv.nleaves, hipMemcpyDeviceToHost));
    AFMM_CHECK(hipMemcpy(scount.data(), tv.seg_count, sizeof(int) * tv.nb, hipMemcpyDeviceToHost));
    AFMM_CHECK(hipMemcpy(uoff.data(), tv.u_off, sizeof(int) * (tv.nleaves + 1), hipMemcpyDeviceToHost));
    if (tv.n_u) AFMM_CHECK(hipMemcpy(uidx.data(), tv.u_idx, sizeof(int) * tv.n_u, hipMemcpyDeviceToHost));
    for (int b = 0; b < tv.nb; ++b) st.n_colleague += nc[b], st.n_v += nv[b];
    int64_t pairs = 0;
    for (int li = 0; li < tv.nleaves; ++li) {
        const int nt = scount[lid[li]];
        st.max_leaf_occ = std::max(st.max_leaf_occ, nt);
        int64_t srcs = 0;
        for (int e = uoff[li]; e < uoff[li + 1]; ++e) srcs += scount[uidx[e]];
        pairs += (int64_t)nt * srcs;
    }
    st.p2p_pairs = pairs - n;
    st.ms_tree = elapsed(c->ev[0], c->ev[1]);
    st.ms_upward = elapsed(c->ev[2], c->ev[3]);
    st.ms_lists = elapsed(c->ev[3], c->ev[7]);  // M2L ordering / source table
    st.ms_m2l = elapsed(c->ev[7], c->ev[4]);
    st.ms_downward = elapsed(c->ev[4], c->ev[5]);
    st.ms_leaf = elapsed(c->ev[5], c->ev[6]);
    st.ms_total = elapsed(c->ev[0], c->ev[6]);
    return AFMM_OK;
}

}  // namespace

extern "C" {

const char* afmm_strerror(int code) {
    switch (code) {
        case AFMM_OK: return "ok";
        case AFMM_ERR_ARG: return "invalid argument";
        case AFMM_ERR_ORDER: return "expansion order outside [1, 12]";
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
        AFMM_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (auto& e : c->ev) AFMM_CHECK(hipEventCreate(&e));
        const Tables& T = tables_for_order(prm->p);
        const std::vector<double> m2m = mfma_a_layout(T.m2m, T.S, 8);
        const std::vector<double> l2l = mfma_a_layout(T.l2l, T.S, 8);
        const std::vector<double> m2l = mfma_a_layout(T.m2l, T.S, kNumOffsets);
        c->d_m2m_sw = c->pool.get<double>("T_m2m_sw", m2m.size());
        c->d_l2l_sw = c->pool.get<double>("T_l2l_sw", l2l.size());
        c->d_m2l_sw = c->pool.get<double>("T_m2l_sw", m2l.size());
        AFMM_CHECK(hipMemcpy(c->d_m2m_sw, m2m.data(), sizeof(double) * m2m.size(), hipMemcpyHostToDevice));
        AFMM_CHECK(hipMemcpy(c->d_l2l_sw, l2l.data(), sizeof(double) * l2l.size(), hipMemcpyHostToDevice));
        AFMM_CHECK(hipMemcpy(c->d_m2l_sw, m2l.data(), sizeof(double) * m2l.size(), hipMemcpyHostToDevice));
        return AFMM_OK;
    });
    if (r) {
        afmm_destroy(c);
        return r;
    }
    *out = c;
    return AFMM_OK;
}

void afmm_destroy(afmm_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->pool.release();
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
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

int afmm_eval_device(afmm_ctx* c, int64_t n, const double* pos, const double* mass, double* acc) {
    if (!c || n < 0 || (n > 0 && (!pos || !mass || !acc))) return AFMM_ERR_ARG;
    if (n >= (1LL << 30)) return fail(c, AFMM_ERR_TOO_LARGE, "n beyond 2^30");
    return guarded(c, [&]() -> int {
        AFMM_CHECK(hipSetDevice(c->prm.device));
        return eval_device(c, n, pos, mass, acc);
    });
}

int afmm_eval(afmm_ctx* c, int64_t n, const double* pos, const double* mass, double* acc) {
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
        AFMM_CHECK(hipMemcpyAsync(dpos, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
        AFMM_CHECK(hipMemcpyAsync(dmass, mass, sizeof(double) * n, hipMemcpyHostToDevice, s));
        const int rc = eval_device(c, n, dpos, dmass, dacc);
        if (rc) return rc;
        AFMM_CHECK(hipMemcpyAsync(acc, dacc, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
        AFMM_CHECK(hipStreamSynchronize(s));
        return AFMM_OK;
    });
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

}  // extern "C"
