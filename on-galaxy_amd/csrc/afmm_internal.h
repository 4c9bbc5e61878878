// ============================================================================
// afmm_internal.h -- device-side data layout and the host entry points of the
// kernel files (scan_sort.hip, tree.hip, fmm.hip). Not part of the C ABI.
//
// HBM layout (all struct-of-arrays, grow-only pools owned by the context):
//   bodies     N: sorted x, y, z, m (FP64), orig index (int32), leaf of body.
//              Input AoS positions are read once by the key kernel and once
//              by the gather.
//   boxes      B, breadth-first: level-major, Morton order inside a level, so
//              the children of a box are contiguous (child_first, child mask).
//              seg_first/seg_count index the sorted bodies.
//   lists      colleague (27 slots/box) and V (189 slots/box); U, W (per
//              leaf) and X (per box) in CSR. M2L builds its own transposed
//              source table srcT[343][targets in (level, octant) order].
//   expansions Mhat[B][S], Lhat[B][S] (FP64, scaled real basis).
// ============================================================================
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/afmm.h"

#include <cstdint>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace afmm {

constexpr int kWave = 64;
constexpr int kMaxDepth = 24;   // AdaptiveOctree::maxDepth (AdaptiveOctree.hpp:49)
constexpr int kCollSlots = 27;  // same-level neighbours incl. self
constexpr int kVSlots = 189;    // 6^3 - 3^3 separated children of the parent's colleagues

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};
#define AFMM_CHECK(call)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::afmm::HipError(e_, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Grow-only named device allocations (and a few pinned host ones, for small
// read-backs that must not block the host when they are enqueued).
class Pool {
public:
    ~Pool() { release(); }
    void release() {
        for (auto& kv : slots_) (void)hipFree(kv.second.p);
        slots_.clear();
        for (auto& kv : hslots_) (void)hipHostFree(kv.second.p);
        hslots_.clear();
    }
    // Pinned host buffer of at least `count` elements (contents undefined
    // after growth). The caller makes sure no copy into it is still pending.
    template <class T>
    T* host(const std::string& name, size_t count) {
        const size_t bytes = (count ? count : 1) * sizeof(T);
        Slot& sl = hslots_[name];
        if (sl.bytes < bytes) {
            void* np = nullptr;
            hipError_t e = hipHostMalloc(&np, bytes, hipHostMallocDefault);
            if (e != hipSuccess) throw HipError(e, "hipHostMalloc(" + name + ")");
            if (sl.p) AFMM_CHECK(hipHostFree(sl.p));
            sl.p = np;
            sl.bytes = bytes;
        }
        return static_cast<T*>(sl.p);
    }
    // Buffer of at least `count` elements; contents are undefined after growth
    // unless keep is set (then the old prefix is copied on `s`).
    template <class T>
    T* get(const std::string& name, size_t count, hipStream_t s = nullptr, bool keep = false) {
        const size_t bytes = (count ? count : 1) * sizeof(T);
        Slot& sl = slots_[name];
        if (sl.bytes < bytes) {
            size_t nb = bytes + bytes / 4 + 256;
            void* np = nullptr;
            hipError_t e = hipMalloc(&np, nb);
            if (e != hipSuccess) throw HipError(e, "hipMalloc(" + name + ")");
            if (keep && sl.p) AFMM_CHECK(hipMemcpyAsync(np, sl.p, sl.bytes, hipMemcpyDeviceToDevice, s));
            if (sl.p) {
                if (keep) AFMM_CHECK(hipStreamSynchronize(s));
                AFMM_CHECK(hipFree(sl.p));
            }
            sl.p = np;
            sl.bytes = nb;
        }
        return static_cast<T*>(sl.p);
    }
    size_t total_bytes() const {
        size_t t = 0;
        for (const auto& kv : slots_) t += kv.second.bytes;
        return t;
    }

private:
    struct Slot {
        void* p = nullptr;
        size_t bytes = 0;
    };
    std::map<std::string, Slot> slots_;
    std::map<std::string, Slot> hslots_;
};

// Root cube and sizing, computed on the device from the bodies.
struct RootInfo {
    double c[3];      // 0.5*(lo+hi) per axis (AdaptiveOctree.hpp:77)
    double half;      // after *(1+1e-9)      (:78-80)
    double tol;       // 1e-9*half            (:81)
    int key_levels;   // octant digits needed: levels that may split under the eps floor
    int nonfinite;    // count of NaN/Inf inputs
};

// Device pointers of the built tree (breadth-first box ids).
struct TreeView {
    int n = 0, nb = 0, nleaves = 0, max_level = 0;
    int nsys = 1;                // systems (roots = boxes 0..nsys-1)
    double tol = 0;              // tolerance of system 0
    const RootInfo* roots;       // [nsys]
    int* bsys;                   // [B] system of each box
    const double* sys_G;         // [nsys] gravitational constant per system
    const double* sys_eps2;      // [nsys] softening^2 per system
    double *x, *y, *z, *m;
    int* orig;
    int* leaf_of;        // [n] BFS leaf id of each sorted body
    int *seg_first, *seg_count, *level, *parent, *octant, *child_first, *nchild;
    unsigned char* child_mask;
    int* refid;          // [B] reference (DFS-allocation) box id
    double* cfp;         // [3B] reference floating-point centres
    double* cdd;         // [6B] exact centres, (hi, lo) per axis
    double* half;        // [B]
    double* invh;        // [B]
    int* ic;             // [3B] integer coordinates at the box's level
    int* leaf_ids;       // [nleaves] BFS ids of the leaves
    int* lo_order;       // [B] boxes sorted by (level, octant); level l is [level_off[l], level_off[l+1])
    int* m2l_order;      // [B] boxes sorted by (octant, level): the M2L target order
    // M2L target tiles over m2l_order, cut inside octant groups: tile k is
    // ranks [m2l_tiles{256,128}[k], [k+1]), for workgroups of at most 256
    // (128) targets; below a CU's worth of full tiles, tiles are cut smaller
    // (a multiple of 64 targets, or 16 or 32 for the 256 plan) so that they
    // spread over more CUs
    int *m2l_tiles256, *m2l_tiles128;
    int m2l_ntiles256 = 0, m2l_ntiles128 = 0;
    int m2l_t256 = 256;  // targets per tile of the 256 plan (16 and 32: row-split workgroups)
    int* nonleaf_ids;    // [B - nleaves] non-leaf boxes in BFS order
    // lists
    int *coll, *ncoll;   // [27B], [B]
    int *vl, *nv;        // [189B], [B]
    int *u_off, *u_idx;  // per leaf (index into leaf_ids)
    int *w_off, *w_idx;  // per leaf
    int *x_off, *x_idx;  // per box
    int64_t n_u = 0, n_w = 0, n_x = 0;
    // device totals of the lists: {colleague entries, V entries, largest leaf,
    // sum over leaves of n_leaf * (bodies in its U leaves)}
    unsigned long long* list_totals;  // [4]
};

// Host-side shape of the built tree: boxes of level l are the breadth-first
// ids [level_off[l], level_off[l+1]).
struct TreeMeta {
    std::vector<int> level_off;
    std::vector<int> nonleaf_off;  // non-leaf boxes of level l: nonleaf_ids[nonleaf_off[l] .. nonleaf_off[l+1])
};

// ---- scan_sort.hip -----------------------------------------------------------
int scan_tmp_ints(int n);
// Exclusive prefix sum of n int32 values (in place allowed); total to *d_total
// when non-null.
void scan_exclusive(const int* d_in, int* d_out, int n, int* d_total, int* d_tmp, hipStream_t s);
// The same over a count held on the device: rng[1] - rng[0] items (at most
// nmax); writes rng[1] + total to *d_end. d_tmp needs scan_tmp_ints(nmax).
void scan_exclusive_range(const int* d_in, int* d_out, const int* d_rng, int nmax, int* d_end, int* d_tmp,
                          hipStream_t s);

struct SortBufs {
    uint64_t* khi[2];
    uint32_t* klo[2];
    int* val[2];
    int* hist;
    int* scan_tmp;
};
int sort_blocks(int n);
// Stable LSD radix sort of (khi, klo, val) on 9-bit digits: digit 0 is klo
// bits [0, 9), digits 1..7 are khi bits [9(g-1), 9g), digits 8 and 9 are klo
// bits [9, 18) and [18, 27). Sorts digits first..first+count-1 in that order,
// starting from buffer set `start`; returns the buffer set holding the result.
// lo = false: the klo words are all zero and stay where they are (digits 1..7
// only; the result's klo is then the input's).
int radix_sort(SortBufs& b, int n, int first_digit, int ndigits, hipStream_t s, int start = 0, bool lo = true);

// ---- tree.hip ----------------------------------------------------------------
// Builds the adaptive octree and all interaction lists for n bodies given as
// device AoS positions and masses. Reproduces AdaptiveOctree::build and
// buildLists (AdaptiveOctree.hpp:65-189) exactly.
// A forest of adaptive octrees, one per system: system q owns the bodies
// [h_sys_off[q], h_sys_off[q+1]) (at least one each), has its own root cube,
// tolerance and leaf-edge floor h_eps_floor[q]; box q of level 0 is its root.
// Each tree, its box numbering (refid) and its lists are those the reference
// builds for that system alone. roots[q].nonfinite > 0 aborts the build.
// on_core (optional) is called once the boxes, bodies, geometry and level
// orders are enqueued and before the interaction lists are: everything the
// upward pass reads is then in place on s. on_vlists (optional) is called
// once the colleague and V lists are enqueued, before U, W and X. With ls
// (and its event ls_ready), the U, W and X lists are built on stream ls after
// it waits for the colleague lists on s (recorded before on_vlists runs; their
// size readback then synchronises ls only, so work that on_vlists enqueues on s
// keeps the GPU busy meanwhile). levels_hint (optional, in and out): the key
// depth of the previous build, taken as a guess that saves the root cubes'
// own round trip (verified, and the sort redone if too shallow); set to this
// build's depth. depth_hint (optional, in and out): the previous tree's
// deepest level, taken as a guess of the levels the key sort must order
// (verified by the topology passes, all levels sorted again if too shallow);
// set to this tree's deepest level.
using TreeHook = std::function<void(const TreeView&, const TreeMeta&)>;
void build_forest(Pool& pool, hipStream_t s, const double* d_pos, const double* d_mass, int nsys,
                  const int* h_sys_off, const double* h_eps_floor, int cap, TreeView& tv, TreeMeta& meta,
                  std::vector<RootInfo>& roots, const TreeHook& on_core = {}, const TreeHook& on_vlists = {},
                  hipStream_t ls = nullptr, hipEvent_t ls_ready = nullptr, int* levels_hint = nullptr,
                  int* depth_hint = nullptr);

// ---- fmm.hip -----------------------------------------------------------------
struct FmmTimes {
    float upward = 0, m2l = 0, downward = 0, leaf = 0;
};
struct FmmTables {
    const double *m2m_sw, *l2l_sw;  // device, 8 octant operators in the MFMA A-fragment layout
    const double* m2l_sw;           // device, 343 offset operators in the same layout
    const float2* m2l_hl;           // device, the same as FP32 (hi, lo) pairs, or null
    const uint4* m2l_bf;            // device, three-piece bf16 layout (mfma_bf16_layout), or null
};
// precision of an evaluation (the values of AFMM_PRECISION_*)
constexpr int kPrecFP64 = 0, kPrecMixed = 1, kPrecMixedBf16 = 2, kPrecM2LBf16 = 3, kPrecP2PFp32 = 4;
// The M2L form of a precision mode (kPrecFP64, kPrecMixed: split FP32, or
// kPrecMixedBf16: three-piece bf16) and whether its P2P runs in FP32.
inline int m2l_form(int prec) {
    return prec == kPrecMixed ? kPrecMixed : (prec == kPrecMixedBf16 || prec == kPrecM2LBf16) ? kPrecMixedBf16 : kPrecFP64;
}
inline bool p2p_fp32(int prec) { return prec == kPrecMixed || prec == kPrecMixedBf16 || prec == kPrecP2PFp32; }
// Runs the FMM passes on a built tree (forest) and writes G*(far + near) to
// d_acc in input order (AoS n x 3) and, when d_phi is non-null, the potential
// -G*(far + near) per body, with G and eps of each body's system (tv.sys_G,
// tv.sys_eps2). prec kPrecMixed: M2L as split-FP32 MFMA; kPrecMixedBf16: M2L
// as three-piece bf16 MFMA; both with P2P in FP32 and FP64 accumulation;
// everything else FP64.
// ev: 8 events or null. reserve_only: take every buffer from the pool, launch nothing.
// side (optional): a second stream with two events; the P2P near field then
// runs on it concurrently with the upward pass, M2L, downward pass, L2P and
// M2P on s, and s joins it before the final combine. Without it (e.g. under
// stream capture) P2P runs on s and combines in the same kernel; both give
// the same bits.
struct SideStream {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // set when run_upward already enqueued P2M/M2M on s (during the list
    // build); up_done is recorded after them and s is waited on before M2L
    bool upward_launched = false;
    hipEvent_t up_done = nullptr;
    // set when run_m2l_prep already enqueued the M2L preparation on the
    // evaluation's own stream (after the V lists)
    bool prep_done = false;
    // set when run_m2l already enqueued M2L itself there too (the U, W and X
    // lists then build on the second stream beside it)
    bool m2l_launched = false;
    // set when run_p2p_near already enqueued P2P on the second stream (after
    // the lists there); lists_done is recorded on it between the lists and P2P,
    // and s waits for it before P2L and M2P
    bool p2p_launched = false;
    hipEvent_t lists_done = nullptr;
    // set when run_far_passes already enqueued P2L (into Lx, then x_done) and
    // M2P (into farw, then w_done) on the second stream
    bool far_beside = false;
    hipEvent_t x_done = nullptr, w_done = nullptr;
};
// The M2L preparation alone on stream s (needs the V lists and the M2L order);
// a run_fmm with side.prep_done then skips it.
// Kernel shapes (afmm_shapes of afmm.h with the defaults resolved): waves per
// M2L workgroup (16 targets each), the leaf size above which a P2P lane takes
// two targets, the fewest lanes per P2P target, and the P2P schedule (0: leaves
// in XCD chunks in tree order; 1: persistent workgroups over the leaves sorted
// by work bin, largest first).
struct Shapes {
    int m2l_waves = 16;  // 8 or 16
    int p2p_pair_bodies = 32;
    int p2p_min_group = 8;
    int p2p_schedule = 0;
};
// The default M2L workgroup of order p and precision mode prec, from
// measurements on the 1M merger: FP64 16 waves up to p = 8 (one wave per
// 16-target block) and from p = 14 (two waves per block, half the row blocks
// each: p = 14 48.8 -> 41.9 ms, p = 16 78.7 -> 73.8 ms per evaluation, where
// 8 waves spilled), 8 waves at p = 9..13 (p = 10 and 12 8% and 1% faster than
// the 16-wave split); split FP32 8 waves (7.6 against 8.5 ms at p = 8); the
// bf16 form 16 waves up to p = 7, 8 above (its FP64 fold totals).
inline int default_m2l_waves(int p, int prec) {
    const int form = m2l_form(prec);
    if (form == kPrecMixedBf16) return p <= 7 ? 16 : 8;
    if (form == kPrecMixed) return 8;
    return p <= 8 || p >= 14 ? 16 : 8;
}
void run_m2l_prep(Pool& pool, hipStream_t s, const TreeView& tv, int p, int prec, const Shapes& sh);
// P2M and M2M alone on stream s (events e0/e1 around them when non-null); a
// run_fmm with side.upward_launched then skips them.
void run_upward(Pool& pool, hipStream_t s, const TreeView& tv, const TreeMeta& meta, int p, const FmmTables& tab,
                hipEvent_t e0, hipEvent_t e1);
// M2L alone on stream s, after the M2L preparation (the split of the
// multipoles into bf16 planes included); events e0/e1 around it when non-null.
void run_m2l(Pool& pool, hipStream_t s, const TreeView& tv, int p, const FmmTables& tab, int prec, const Shapes& sh,
             hipEvent_t e0, hipEvent_t e1);
// The P2P near field alone on stream s into the pool's near-field buffers
// (potential too when pot), for a later run_fmm with side.p2p_launched.
void run_p2p_near(Pool& pool, hipStream_t s, const TreeView& tv, int prec, const Shapes& sh, bool pot, hipEvent_t e0,
                  hipEvent_t e1);
// P2L over X into the pool's Lx and M2P over W into its farw (and farwpot when
// pot) on stream s, recording x_done and w_done after each, for a later
// run_fmm with side.far_beside.
void run_far_passes(Pool& pool, hipStream_t s, const TreeView& tv, int p, bool pot, hipEvent_t x_done,
                    hipEvent_t w_done);
void run_fmm(Pool& pool, hipStream_t s, const TreeView& tv, const TreeMeta& meta, int p,
             const FmmTables& tab, int prec, const Shapes& sh, double* d_acc, double* d_phi,
             hipEvent_t* ev, bool reserve_only = false, SideStream side = SideStream{});

// ---- integrate.hip -----------------------------------------------------------
// One system of a leapfrog step: its resident arrays and step settings.
struct JobRef {
    double *pos, *vel, *acc;  // [3n] AoS, input order
    double* phi;              // [n] or null
    const double* mass;       // [n]
    double dt;
    int init;   // accelerations hold the forces of the current positions
    int drift;  // 1: drift before and kick after the evaluation; 0: evaluate only
};
// Batch body k is body k - boff[q] of system q (boff[q] <= k < boff[q+1]).
void leapfrog_gather(hipStream_t s, int n, int njobs, const int* d_boff, const JobRef* d_jobs, double* bpos,
                     double* bmass);
void leapfrog_kick(hipStream_t s, int n, int njobs, const int* d_boff, const JobRef* d_jobs, const double* bacc,
                   const double* bphi);
// d_out[9q..9q+9) = {KE, PE = 0.5 sum m phi, KE+PE, momentum xyz, angular momentum xyz}
void diagnostics(hipStream_t s, int njobs, const JobRef* d_jobs, const int* d_n, double* d_out);

// ---- host_io.hip -------------------------------------------------------------
// Host <-> device copies of the drop-in entry points through two pinned
// staging buffers per context: the host copies one chunk (several threads)
// while the DMA of the previous one runs. upload returns once the host array
// has been read (the DMA may still run on s); download returns with the host
// array written.
class HostStager {
public:
    HostStager() = default;
    HostStager(const HostStager&) = delete;
    HostStager& operator=(const HostStager&) = delete;
    ~HostStager();
    void upload(void* d_dst, const void* h_src, size_t bytes, hipStream_t s);
    void download(void* h_dst, const void* d_src, size_t bytes, hipStream_t s);

private:
    void reserve();
    void* pin_[2] = {nullptr, nullptr};
    hipEvent_t ev_[2] = {nullptr, nullptr};
};

// ---- direct.hip --------------------------------------------------------------
// FP64 direct sum of the n bodies (pos AoS n x 3) onto the bodies d_tgt[0..ntgt):
// acc (AoS ntgt x 3) and, if d_phi is non-null, the softened potential.
void direct_targets(Pool& pool, hipStream_t s, const double* d_pos, const double* d_mass, int n, const int* d_tgt,
                    int ntgt, double G, double eps, double* d_acc, double* d_phi);
// The same sum with its partial sums in a static device array instead of the
// pool: no allocation. Targets are device int64 indices (outside [0, n) give
// NaN); at most kNoAllocMaxTargets of them. direct_scratch_setup(device) must
// have run once with that device current (afmm_create does it).
constexpr int kMaxDevices = 64;
constexpr int kNoAllocMaxTargets = AFMM_DIRECT_NOALLOC_MAX_TARGETS;
void direct_scratch_setup(int device);
void direct_targets_noalloc(int device, hipStream_t s, const double* d_pos, const double* d_mass, int n,
                            const int64_t* d_tgt, int ntgt, double G, double eps, double* d_acc, double* d_phi);

}  // namespace afmm
