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
/* FP64: every phase in FP64 (the reference CPU path's arithmetic).
 * MIXED: M2L as split-FP32 MFMA products (each operand carried as an FP32
 * hi/lo pair, three products per block) and P2P in FP32 with FP64
 * accumulation per target; tree, P2M/M2M/L2L/P2L, L2P/M2P and the pair
 * differences stay FP64. The CUDA twin's reduced precision is its `real`
 * template (gpu/AdaptiveFMM3DGPU.cu:219-334). */
/* MIXED_BF16: as MIXED, but M2L carries each operand as three bf16 pieces
 * (hi + mid + lo, ~2^-24 relative) and takes the six products at or above
 * that order on v_mfma_f32_16x16x32_bf16, FP32 accumulation: the same
 * compensated split on the matrix type MI355X runs 16x faster than FP32. */
#define AFMM_PRECISION_FP64 0
#define AFMM_PRECISION_MIXED 1
#define AFMM_PRECISION_MIXED_BF16 2
/* Attribution modes (no counterpart in the reference): one of the two
 * reduced-precision phases alone, everything else FP64, so that the error
 * each contributes can be stated apart. M2L_BF16: the bf16 M2L of
 * MIXED_BF16 with the FP64 P2P; P2P_FP32: the FP64 M2L with the FP32 P2P of
 * the mixed modes. */
#define AFMM_PRECISION_M2L_BF16 3
#define AFMM_PRECISION_P2P_FP32 4

/* Solver parameters: AdaptiveFMM3D(KernelParams3D{G, eps}, p, threshold)
 * (AdaptiveFMM3D.hpp:22-23; Kernel3D.hpp:24-27). The eps also floors the leaf
 * edge, as the reference passes params.eps to AdaptiveOctree::build
 * (AdaptiveFMM3D.hpp:37). */
typedef struct {
    double G;
    double eps;
    int32_t p;          /* expansion order, 1..AFMM_MAX_ORDER */
    int32_t leaf_cap;   /* split a box while it holds > leaf_cap bodies (64) */
    int32_t precision;  /* AFMM_PRECISION_* */
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
    double ms_p2p; /* the P2P near-field kernel alone; outside graph mode it runs on a second
                      stream beside the upward pass through M2P */
    /* graph mode (afmm_set_graph): executable graphs built and updated in
     * place so far; in graph mode only ms_tree and ms_total are timed */
    int64_t graph_instantiations, graph_updates;
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

/* Batched evaluation of nsys independent systems with one launch per FMM
 * phase over all of them (a forest of octrees, one root per system). System q
 * owns bodies [sys_off[q], sys_off[q+1]) of the concatenated arrays
 * (sys_off[0] = 0, strictly increasing: every system has a body) and uses
 * G[q], eps[q] (either array may be NULL: the context's values). Expansion
 * order, leaf capacity and precision are the context's. Each system's tree,
 * accelerations and potentials are those of afmm_eval on that system alone,
 * bit for bit. Host arrays (afmm_eval_batch) or device arrays
 * (afmm_eval_batch_device; sys_off, G and eps stay host arrays); phi may be
 * NULL. There is no reference counterpart: the reference evaluates one
 * system per call (AdaptiveFMM3D.hpp:29-110). */
int afmm_eval_batch(afmm_ctx* ctx, int32_t nsys, const int64_t* sys_off, const double* G, const double* eps,
                    const double* pos_xyz, const double* mass, double* acc_xyz, double* phi);
int afmm_eval_batch_device(afmm_ctx* ctx, int32_t nsys, const int64_t* sys_off, const double* G,
                           const double* eps, const double* d_pos_xyz, const double* d_mass, double* d_acc_xyz,
                           double* d_phi);

/* ---- device-resident time integration ----------------------------------- */
/* Leapfrog state of nsys independent systems kept in HBM between steps
 * (bodies, velocities, accelerations), integrated with the reference's
 * velocity Verlet (VelocityVerlet3D, Integrator3D.hpp:19-40):
 *   x += v*dt + a*(0.5*dt*dt);  a' = forces(x);  v += (a + a')*(0.5*dt);  a = a'
 * in the reference's operation order per component, with one batched force
 * evaluation (afmm_eval_batch semantics: sys_off, per-system G and eps) per
 * step for all systems. The reference round-trips host arrays through its
 * solver every step (AdaptiveFMM3DGPU.cu:252-307). */
typedef struct afmm_state afmm_state;

/* Uploads the systems and computes their initial accelerations
 * (VelocityVerlet3D::initialize). The state uses ctx (order, leaf capacity,
 * precision, stream) for its evaluations; destroy it before ctx. */
int afmm_state_create(afmm_ctx* ctx, int32_t nsys, const int64_t* sys_off, const double* G, const double* eps,
                      const double* pos_xyz, const double* vel_xyz, const double* mass, afmm_state** out);
/* nsteps velocity-Verlet steps of every system with time step dt. */
int afmm_state_step(afmm_state* st, double dt, int32_t nsteps);
/* Copies the current positions, velocities and accelerations to host arrays
 * (input order; any pointer may be NULL). */
int afmm_state_download(afmm_state* st, double* pos_xyz, double* vel_xyz, double* acc_xyz);
/* Diagnostics3D per system (Diagnostics3D.hpp:20-38) into out[9*nsys]:
 * {kinetic, potential, total, momentum xyz, angular momentum xyz}. The
 * potential energy is 0.5 sum_i m_i phi_i with the FMM potential of
 * afmm_eval_potential (one evaluation), not the reference's O(N^2) pair sum;
 * the reductions are FP64 in a fixed order. */
int afmm_state_diagnostics(afmm_state* st, double* out);
void afmm_state_destroy(afmm_state* st);

/* ---- request server ------------------------------------------------------- */
/* Continuous batching of independent jobs on one GPU. A job is a system
 * (n bodies, G, eps, order p) and either one force evaluation (steps = 0) or
 * `steps` velocity-Verlet steps of size dt (afmm_state semantics). Jobs are
 * admitted in arrival order while the sum of their footprint estimates
 * (afmm_job_bytes) fits hbm_budget; admitted jobs of the same order are
 * stepped together, one batched evaluation per step (afmm_eval_batch), and
 * jobs that arrive meanwhile join at the next step. Each order runs on its
 * own worker thread and stream, so batches of different orders overlap. A
 * job's result is bit-identical to running it alone. The reference has no
 * counterpart (its drivers step one system per call, gpu/galaxy_sim.cu:230-237). */
typedef struct afmm_server afmm_server;
typedef struct {
    int64_t hbm_budget;  /* bytes of device memory jobs may claim (<= 0: 240 GiB) */
    int32_t leaf_cap;    /* leaf capacity of every job (64) */
    int32_t precision;   /* AFMM_PRECISION_FP64, _MIXED or _MIXED_BF16 */
    int32_t device;      /* HIP device ordinal */
    int32_t max_batch;   /* most jobs in one batched step of an order (<= 0: no limit) */
} afmm_server_params;
typedef struct {
    int64_t n;
    const double* pos_xyz;  /* copied at submission */
    const double* vel_xyz;  /* may be NULL when steps == 0 */
    const double* mass;
    double G, eps, dt;
    int32_t p;              /* expansion order 1..AFMM_MAX_ORDER */
    int32_t steps;          /* 0: force evaluation; k > 0: k leapfrog steps */
} afmm_job_desc;
#define AFMM_JOB_QUEUED 0
#define AFMM_JOB_RUNNING 1
#define AFMM_JOB_DONE 2
#define AFMM_JOB_FAILED 3
typedef struct {
    int32_t state;       /* AFMM_JOB_* */
    int32_t error;       /* afmm_status of a failed job */
    int64_t hbm_bytes;   /* admission estimate */
    double t_submit, t_admit, t_done;  /* seconds since afmm_server_create */
    int32_t steps_done;  /* leapfrog steps completed */
    int32_t batches;     /* batched evaluations the job took part in */
} afmm_job_info;

/* Footprint estimate used for admission: resident arrays, the job's share of
 * the batch, tree and sort buffers, its expansions and lists, at leaf
 * capacity 64 (a server scales the box count by its own leaf_cap). */
int64_t afmm_job_bytes(int64_t n, int32_t p);
int afmm_server_create(const afmm_server_params* prm, afmm_server** out);
/* Prepares the worker of order p (context, translation tables) ahead of jobs. */
int afmm_server_warm(afmm_server* sv, int32_t p);
int afmm_server_submit(afmm_server* sv, const afmm_job_desc* job, int64_t* id);
/* Blocks until the job is done or failed; returns its afmm_status. */
int afmm_server_wait(afmm_server* sv, int64_t id);
/* Blocks until every submitted job is done or failed. */
int afmm_server_drain(afmm_server* sv);
int afmm_server_info(afmm_server* sv, int64_t id, afmm_job_info* info);
/* Final positions, velocities and accelerations of a done job (any pointer
 * may be NULL); the job is then forgotten. */
int afmm_server_result(afmm_server* sv, int64_t id, double* pos_xyz, double* vel_xyz, double* acc_xyz);
/* Sum of the footprint estimates of the admitted, unfinished jobs. */
int64_t afmm_server_hbm_in_use(afmm_server* sv);
/* Finishes running jobs, fails queued ones (AFMM_ERR_STATE; their waiters
 * return), joins the workers. A thread blocked in afmm_server_wait or
 * afmm_server_drain must return before destroy frees the server: callers do
 * not let those calls race destroy. */
void afmm_server_destroy(afmm_server* sv);

/* FP64 direct summation over all n bodies onto the targets tgt[0..ntgt)
 * (indices into the bodies), with the context's G and eps: the accuracy
 * yardstick DirectSolver3D (DirectSolver3D.hpp:15-34) restricted to a target
 * subsample, and the softened potential of Kernel3D.hpp:39-42:
 *   acc_xyz[k] = G sum_{j != tgt[k]} m_j d / (|d|^2 + eps^2)^{3/2}
 *   phi[k]     = -G sum_{j != tgt[k]} m_j / sqrt(|d|^2 + eps^2)   (phi may be NULL)
 * Host arrays; the sum order is fixed, so results are reproducible. */
int afmm_direct(afmm_ctx* ctx, int64_t n, const double* pos_xyz, const double* mass, int64_t ntgt,
                const int64_t* tgt, double* acc_xyz, double* phi);

/* Allocation-free afmm_direct on device arrays. It has no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops). Same sums, same bits as afmm_direct
 * for the same bodies and targets: d_tgt holds ntgt int64 body indices on the
 * device (an index outside [0, n) gives NaN for that target; nothing outside
 * the arrays is read). The only scratch, the per-split partial sums, is a
 * static device array of 10 MiB that each GPU the library runs on holds until
 * the process ends, so ntgt <= AFMM_DIRECT_NOALLOC_MAX_TARGETS
 * (AFMM_ERR_TOO_LARGE beyond). Work is enqueued on `stream` (a hipStream_t;
 * NULL: the context's stream) and not waited for. Every call on a GPU shares
 * that GPU's scratch: calls that overlap on one GPU race. */
#define AFMM_DIRECT_NOALLOC_MAX_TARGETS 65536
int afmm_direct_device_noalloc(afmm_ctx* ctx, int64_t n, const double* d_pos_xyz, const double* d_mass, int64_t ntgt,
                               const int64_t* d_tgt, double* d_acc_xyz, double* d_phi, void* stream);

/* afmm_eval / afmm_eval_device plus the FMM potential per body (input
 * order) when phi is non-null: phi[i] = -G (far-field expansions at x_i +
 * sum_{j in U, j != i} m_j / sqrt(|d|^2 + eps^2)), the approximation of
 * -G sum_{j != i} m_j / sqrt(|d|^2 + eps^2) (Kernel3D.hpp:39-42) whose
 * gradient the accelerations are; softening acts in the near field only, as
 * for the accelerations. FP64 in both precision modes except the near-field
 * sum, which follows the context's precision. 0.5 sum_i m_i phi_i is the
 * potential energy of Diagnostics3D (Diagnostics3D.hpp:20-38). */
int afmm_eval_potential(afmm_ctx* ctx, int64_t n, const double* pos_xyz, const double* mass, double* acc_xyz,
                        double* phi);
int afmm_eval_device_potential(afmm_ctx* ctx, int64_t n, const double* d_pos_xyz, const double* d_mass,
                               double* d_acc_xyz, double* d_phi);

/* Block until the context's stream is idle. */
int afmm_synchronize(afmm_ctx* ctx);

/* Graph mode (no counterpart in the reference, which launches from the host
 * per pass): when enabled, each evaluation captures its FMM passes (M2L
 * through P2P, or P2M through P2P with afmm_set_overlap off; after the tree
 * build, which needs host round trips) into a HIP graph and launches that.
 * With the overlap on, the P2P near field is a branch of the graph on the
 * context's own second stream, beside M2L and the passes after it. An
 * evaluation whose passes have the same shape as the previous one updates the
 * executable graph in place. Results are the same as without it. */
int afmm_set_graph(afmm_ctx* ctx, int enable);

/* Kernel shapes (no counterpart in the reference, whose CUDA twin runs one
 * thread per target box or body): chosen by the project's autotuner
 * (on-galaxy_amd/tune.py) or left at their defaults. A zero field keeps the
 * default. Every shape gives the same tree and lists; M2L waves and the P2P
 * schedule also leave the results' bits unchanged, while the P2P lane grouping
 * changes the order of each near-field sum (same accuracy, different last bits). */
typedef struct {
    int32_t m2l_waves;        /* waves per M2L workgroup: 8 or 16. One wave per 16 targets, except
                                 16 waves of the FP64 M2L above p = 8 and of the split-FP32 M2L,
                                 where two waves share 16 targets (half the row blocks each).
                                 Default 16 for FP64 at p <= 8 and p >= 14 and for bf16 at p <= 7,
                                 else 8 */
    int32_t p2p_pair_bodies;  /* leaves with more bodies give each P2P lane two targets (default 32) */
    int32_t p2p_min_group;    /* fewest lanes sharing one target's sources: 1, 2, 4, ..., 64 (default 8) */
    int32_t p2p_schedule;     /* 1: persistent workgroups over the leaves sorted by work bin, largest
                                 first; 2: one workgroup per leaf in XCD-local chunks of tree order
                                 (the default) */
} afmm_shapes;
int afmm_set_shapes(afmm_ctx* ctx, const afmm_shapes* shapes);
/* The shapes in effect, defaults filled in. */
int afmm_get_shapes(const afmm_ctx* ctx, afmm_shapes* out);

/* Two-stream overlap (no counterpart in the reference, which has one host
 * thread and one stream). Modes:
 *   AFMM_OVERLAP_OFF: everything in order on the context's stream
 *     (afmm_stats.ms_p2p then times P2P alone);
 *   AFMM_OVERLAP_AUTO (the default): a second stream, shared by the process's
 *     contexts on that device, runs the upward pass beside the interaction-list
 *     build and P2P beside the far-field passes; with the bf16 or split-FP32
 *     M2L, and the FP64 one up to p = 8, from p = 14 and at p = 9-13 for up to
 *     500,000 bodies, the U/W/X lists, P2L, M2P and P2P
 *     run there beside M2L as well (their VALU work shares the SIMDs with its
 *     MFMAs);
 *   AFMM_OVERLAP_BESIDE_M2L: the latter schedule in every form, with P2P
 *     beside M2L too (the automatic schedule starts P2P after the bf16 M2L);
 *   AFMM_OVERLAP_IN_ORDER: the former (lists, P2L and M2P in order before M2L).
 * In graph mode the captured P2P branch uses the context's own second stream
 * and the lists stay in order. Results are the same bits in every mode. */
#define AFMM_OVERLAP_OFF 0
#define AFMM_OVERLAP_AUTO 1
#define AFMM_OVERLAP_BESIDE_M2L 2
#define AFMM_OVERLAP_IN_ORDER 3
int afmm_set_overlap(afmm_ctx* ctx, int mode);

/* ---- initial conditions (host) -------------------------------------------
 * makePlummerSphere (InitialConditions3D.hpp:100-152): n equal-mass bodies of
 * a Plummer sphere (scale a, total mass total_mass, constant G) drawn from
 * libstdc++'s std::mt19937(seed) through uniform_real_distribution<double>,
 * centre of mass and net momentum removed; the reference's bodies bit for bit
 * for the same seed. Host arrays (pos_xyz, vel_xyz n x 3; mass n); no device
 * work. AFMM_ERR_ARG for n < 1, a <= 0 or non-finite parameters. */
int afmm_make_plummer(int64_t n, double a, double total_mass, double G, uint32_t seed, double* pos_xyz,
                      double* vel_xyz, double* mass);

/* ---- introspection ------------------------------------------------------ */
int afmm_get_stats(const afmm_ctx* ctx, afmm_stats* out);
int afmm_tree_export(afmm_ctx* ctx, afmm_tree* out);

#ifdef __cplusplus
}
#endif
#endif /* AFMM_H */
