/* ============================================================================
 * afmm.h -- C ABI of the MI355X-native 3D adaptive FMM force evaluation.
 *
 * Drop-in boundary for the reference's force-solver seam:
 *   galaxy3d::ForceSolver3D::computeAccelerations      (src3d/ForceSolver3D.hpp:20-28)
 *   galaxy3d::AdaptiveFMM3D(KernelParams3D, p, thr)    (src3d/AdaptiveFMM3D.hpp:22-23, :29-110)
 *   gpu::adaptiveFmm3dGpu<real>(bodies, G, eps, p, thr, GpuTiming*)
 *                                                      (src3d/gpu/AdaptiveFMM3DGPU.cuh:16-18)
 * Plain pointers and sizes only. Every function returns 0 or a negative
 * afmm_status and never exits the process (the reference calls exit(1) on a
 * CUDA error and exit(2) on an unsupported p: AdaptiveFMM3DGPU.cu:23-26, :328).
 * A context is not thread-safe: use one host thread per context.
 * ==========================================================================*/
#ifndef AFMM_H
#define AFMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    AFMM_OK = 0,
    AFMM_ERR_ARG = -1,        /* null pointer, n < 0, bad G/eps */
    AFMM_ERR_ORDER = -2,      /* p outside [1, AFMM_MAX_ORDER] */
    AFMM_ERR_LEAF_CAP = -3,   /* leaf capacity < 1 */
    AFMM_ERR_NONFINITE = -4,  /* a position or mass is NaN or Inf */
    AFMM_ERR_DEVICE = -5,     /* HIP runtime error (see afmm_last_error) */
    AFMM_ERR_NO_MEMORY = -6,  /* device allocation failed */
    AFMM_ERR_STATE = -7,      /* call needs a prior afmm_eval on this context */
    AFMM_ERR_TOO_LARGE = -8,  /* n beyond int32 particle indexing */
    AFMM_ERR_PRECISION = -9   /* precision mode not supported */
} afmm_status;

#define AFMM_MAX_ORDER 16
#define AFMM_PRECISION_FP64 0
#define AFMM_PRECISION_MIXED 1

/* Solver parameters: AdaptiveFMM3D(KernelParams3D{G, eps}, p, threshold)
 * (AdaptiveFMM3D.hpp:22-23; Kernel3D.hpp:24-27). The eps also floors the leaf
 * edge, as the reference passes params.eps to AdaptiveOctree::build
 * (AdaptiveFMM3D.hpp:37). */
typedef struct {
    double G;
    double eps;
    int32_t p;          /* expansion order, 1..AFMM_MAX_ORDER */
    int32_t leaf_cap;   /* split a box while it holds > leaf_cap bodies (64) */
    int32_t precision;  /* AFMM_PRECISION_FP64 */
    int32_t device;     /* HIP device ordinal */
} afmm_params;

typedef struct afmm_ctx afmm_ctx;

/* Sizes of the last evaluation's tree (AdaptiveOctree fields,
 * AdaptiveOctree.hpp:40-47) and the per-phase timings of GpuTiming
 * (gpu/FMM3DGPU.cuh:15-22) plus counts used for roofline accounting. */
typedef struct {
    int64_t n_bodies;
    int64_t n_boxes;
    int64_t n_leaves;
    int64_t n_colleague, n_u, n_v, n_w, n_x; /* total list entries */
    int64_t p2p_pairs;                       /* sum over leaves of n_b * sum_U n_c, minus n */
    int32_t max_level;
    int32_t max_leaf_occ;                    /* AdaptiveOctree::maxLeafOcc (:191-195) */
    double ms_tree, ms_lists, ms_upward, ms_m2l, ms_downward, ms_leaf, ms_total; /* device time */
} afmm_stats;

/* Tree export, in the reference's box numbering and list orders
 * (AdaptiveOctree.hpp:31-47, :137-189). Caller-allocated host buffers sized
 * from afmm_stats; a NULL pointer skips that field. Lists are CSR:
 * *_off has n_boxes+1 entries. */
typedef struct {
    int32_t* parent;      /* [n_boxes] */
    int32_t* child;       /* [8*n_boxes], -1 when absent */
    int32_t* level;       /* [n_boxes] */
    double* center;       /* [3*n_boxes] cx,cy,cz */
    double* half;         /* [n_boxes] */
    int8_t* leaf;         /* [n_boxes] */
    int32_t* pfirst;      /* [n_boxes] */
    int32_t* pcount;      /* [n_boxes] */
    double* px; double* py; double* pz; double* mass; /* [n_bodies] permuted */
    int32_t* orig_index;  /* [n_bodies] permuted slot -> input index */
    int32_t* leaves;      /* [n_leaves] ascending box id */
    int32_t* colleague_off; int32_t* colleague_idx;
    int32_t* u_off; int32_t* u_idx;
    int32_t* v_off; int32_t* v_idx;
    int32_t* w_off; int32_t* w_idx;
    int32_t* x_off; int32_t* x_idx;
    double tol;           /* out: adjacency tolerance 1e-9*half_root */
} afmm_tree;

/* ---- lifecycle ---------------------------------------------------------- */
int afmm_create(const afmm_params* params, afmm_ctx** out);
void afmm_destroy(afmm_ctx* ctx);
const char* afmm_strerror(int code);
const char* afmm_last_error(const afmm_ctx* ctx);

/* ---- force evaluation ----------------------------------------------------
 * AdaptiveFMM3D::computeAccelerations (AdaptiveFMM3D.hpp:29-110):
 * pos_xyz is n x 3 (Vec3 AoS, Types3D.hpp:12-24), mass n, acc_xyz n x 3 in
 * INPUT order, acc = G * (far + near). Host buffers in and out. n == 0
 * returns AFMM_OK with nothing written (AdaptiveFMM3D.hpp:34). */
int afmm_eval(afmm_ctx* ctx, int64_t n, const double* pos_xyz, const double* mass,
              double* acc_xyz);

/* Same evaluation on device-resident buffers (hipMalloc'd by the caller) on
 * the context's stream; no host round-trip of bodies or results. */
int afmm_eval_device(afmm_ctx* ctx, int64_t n, const double* d_pos_xyz, const double* d_mass,
                     double* d_acc_xyz);

/* Block until the context's stream is idle. */
int afmm_synchronize(afmm_ctx* ctx);

/* ---- introspection ------------------------------------------------------ */
int afmm_get_stats(const afmm_ctx* ctx, afmm_stats* out);
int afmm_tree_export(afmm_ctx* ctx, afmm_tree* out);

#ifdef __cplusplus
}
#endif
#endif /* AFMM_H */
