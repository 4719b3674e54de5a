/* fmm_capi.h -- C ABI of the MI355X Biot-Savart fast multipole library.
 *
 * Drop-in boundary for the reference multipole-solver path of iamwhoamianamy/bachelors
 * (FMM/FMMGPU).  One call per reference method:
 *
 *   reference (C++)                                              here
 *   FastMultipoleSolver(quadratures, points, problem, capS, capT)  fmm_create + fmm_set_sources
 *     fast_multipole_solver.hpp:25-30, multipole_solver.cpp:20-36    + fmm_set_targets + fmm_build_trees
 *   MultipoleSolver::calcMultipoleExpansionsAtLeaves()             fmm_leaf_expansions
 *     multipole_solver.hpp:39, multipole_solver.cpp:56-95
 *   MultipoleSolver::calclMultipoleExpansions(M2MAlg, Device)      fmm_upward_with (fmm_upward: Matrices)
 *     multipole_solver.hpp:42-44, multipole_solver.cpp:102-139
 *   FastMultipoleSolver::calcLocalMultipoleExpansions(M2LAlg, Dev) fmm_downward
 *     fast_multipole_solver.hpp:39-41, fast_multipole_solver.cpp:62-102
 *   FastMultipoleSolver::calcB(real current)                       fmm_evaluate_B (+ fmm_leaf_order)
 *     fast_multipole_solver.hpp:46, fast_multipole_solver.cpp:214-231
 *   FastMultipoleSolver::calcA(real current)                       fmm_evaluate_A (+ fmm_leaf_order)
 *     fast_multipole_solver.hpp:45, fast_multipole_solver.cpp:195-212
 *   MultipoleSolver::calcA / calcB(real current, const Vector3&)   fmm_treecode (many points at once)
 *     multipole_solver.hpp:46-47, multipole_solver.cpp:686-700
 *   FastMultipoleSolver(vector<BEMQuadrature>&, ..., Problem::BEM)  fmm_set_bem_sources
 *     fast_multipole_solver.hpp:32-37, multipole_solver.cpp:38-54
 *   FastMultipoleSolver::calcBEM(real current)                     fmm_evaluate_bem (+ fmm_leaf_order)
 *     fast_multipole_solver.hpp:47, fast_multipole_solver.cpp:233-250
 *   MultipoleSolver::getQuadratureOctreeNodeCount()                fmm_tree_size(ctx, 0, ...)
 *     multipole_solver.hpp:40, multipole_solver.cpp:97-100
 *   math::calcBioSavartLaplace(current, point, quadratures)         fmm_direct_B_dev
 *     integration.cpp:101-111
 * and it replaces the reference's device boundary (kernels::translateAll*,
 * translation_algorithms.hpp:13-60; blas::*, blass_callers.hpp:8-65), whose GEMMs now run
 * inside the library on MFMA.  fmm_adjoint_weights_dev has no reference counterpart: it
 * is the transposed FMM giving dL/dw for L = sum_i |B_i - B*_i|^2.
 *
 * Conventions: every function returns FMM_OK (0) or an error code and never exits; the
 * message is in fmm_last_error(ctx).  The context owns all device memory.  Host arrays
 * are copied in; *_dev entry points take device pointers and run on the context's HIP
 * stream, ordered after the caller's work on the default (null) stream by an event (no
 * device-wide synchronisation): work on other caller streams must be finished by the caller,
 * or the caller's stream handed over with fmm_set_stream.  One context per host thread.
 */
#ifndef FMM_CAPI_H
#define FMM_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FMM_OK 0
#define FMM_ERR_ARG 1          /* bad argument */
#define FMM_ERR_STATE 2        /* stage called out of order (reference: "Multipoles are not ready!") */
#define FMM_ERR_HIP 3          /* HIP runtime error */
#define FMM_ERR_LIMIT 4        /* a size limit of the library was exceeded */
#define FMM_ERR_UNSUPPORTED 5  /* option combination not implemented */

#define FMM_PREC_FP64 0        /* translations on FP64 MFMA, FP64 expansions */
#define FMM_PREC_FP32 1        /* FP32 MFMA, FP32 accumulate */
#define FMM_PREC_BF16X3 2      /* BF16 split products hi*hi + hi*lo + lo*hi, FP32 accumulate */
#define FMM_PREC_F16X3 3       /* M2L on FP16 split products (operands equilibrated by powers of 2),
                                  FP32 accumulate; M2M and L2L in FP32 */

#define FMM_FRAME_REFERENCE 0  /* target root = reference bounding box (math.cpp:39-57) */
#define FMM_FRAME_SHARED 1     /* target root = source root: M2L operators shared per offset */

#define FMM_NEAR_NONE 0        /* close lists add nothing */
#define FMM_NEAR_P2P 1         /* close lists summed particle by particle (exact near field) */
#define FMM_NEAR_TREECODE 2    /* the reference: M2P walks from each close node, failing leaves
                                  add nothing (quadrature_octree.cpp:171-258) */

typedef struct fmm_ctx fmm_ctx;

typedef struct fmm_config {
    int order;                    /* p; reference harmonicOrder = 10 (multipole_solver.hpp:25) */
    int source_leaf_capacity;     /* reference quadratureOctreeLeafCapacity (leaf holds <= cap-1) */
    int target_leaf_capacity;     /* reference calculationPointOctreeLeafCapacity (leaf holds <= cap) */
    int precision;                /* FMM_PREC_* */
    int frame;                    /* FMM_FRAME_* */
    int m2l_full;                 /* 1: M2L keeps l+lambda <= 2p; 0: reference l+lambda <= p */
    int max_depth;                /* depth cap (<= 21); the reference has none */
    int near_field;               /* FMM_NEAR_*: P2P (default), none, or the reference's treecode M2P */
    double mac_scale;             /* 1.0 = reference MAC (2Rmax+Rmin)^2 < d^2 (fast_multipole_solver.cpp:167-184) */
    double source_root_center[3]; /* reference Box((0,0,0),(3,3,3)) (multipole_solver.cpp:662-663) */
    double source_root_half[3];
} fmm_config;

typedef struct fmm_stats {
    int source_nodes, target_nodes, source_levels, target_levels;
    int source_slots, target_slots;      /* points kept by each tree */
    int dropped_sources, dropped_targets;
    long long far_pairs, close_pairs, m2l_classes;
    long long near_interactions;         /* sum over close pairs of n_target * n_source */
    int order, padded_length, bfs_steps;
    /* stage times in ms (HIP events on the context stream) */
    double t_tree, t_p2m, t_m2m, t_lists, t_tables, t_m2l, t_l2l, t_l2p, t_p2p, t_total;
    double t_m2l_gemm;                   /* the M2L translation GEMM alone (part of t_m2l) */
    int source_tree_sorted, target_tree_sorted;   /* 1: built by one sort of point paths;
                                                     0: level by level (gaps/overlaps of child boxes) */
    int source_leaves, target_leaves;    /* leaves holding points */
    double t_to_m2l_gemm;                /* evaluation start to the start of the M2L GEMM */
    double t_after_m2l_gemm;             /* end of the M2L GEMM to the end of the evaluation */
    double t_lists_far;                  /* the far pairs and classes (first part of t_lists) */
} fmm_stats;

void fmm_config_init(fmm_config* cfg);
int fmm_create(fmm_ctx** out, const fmm_config* cfg);
void fmm_destroy(fmm_ctx* ctx);
const char* fmm_last_error(const fmm_ctx* ctx);
int fmm_set_stream(fmm_ctx* ctx, void* hip_stream);   /* NULL: the context's own stream */
int fmm_synchronize(fmm_ctx* ctx);

/* inputs: xyz is n x 3 (row-major), w is n, J (n x 3) may be NULL (current along phi-hat,
 * reference integration.cpp:155).  Host and device-pointer variants. */
int fmm_set_sources(fmm_ctx* ctx, size_t n, const double* xyz, const double* w, const double* J);
int fmm_set_targets(fmm_ctx* ctx, size_t m, const double* xyz);
int fmm_set_sources_dev(fmm_ctx* ctx, size_t n, const double* xyz, const double* w, const double* J);
int fmm_set_weights_dev(fmm_ctx* ctx, const double* w);   /* new weights, same sources and trees */
/* Problem::BEM sources (BEMQuadrature, bem_quadrature.hpp; FastMultipoleSolver(vector<BEMQuadrature>&,
 * ...), fast_multipole_solver.hpp:32-37): n points with weight w, field B (n x 3) and normal (n x 3).
 * The reference roots their tree at Box((0,0,0), (1.1, 1.1, 1.6)) (multipole_solver.cpp:668-669):
 * set source_root_center / source_root_half of the config to that box. */
int fmm_set_bem_sources(fmm_ctx* ctx, size_t n, const double* xyz, const double* w, const double* B,
                        const double* normal);
int fmm_set_bem_sources_dev(fmm_ctx* ctx, size_t n, const double* xyz, const double* w, const double* B,
                            const double* normal);
int fmm_set_targets_dev(fmm_ctx* ctx, size_t m, const double* xyz);

/* M2M algorithms of MultipoleSolver::calclMultipoleExpansions (multipole_solver_enums.hpp:3-10,
 * multipole_solver.cpp:102-339); all give the same multipoles.  NO_TRANSLATION: P2M at every node
 * from its points; COMPLEX / REAL / LAYERS: the per-node translation formula without operator
 * tables; MATRICES: the per-(level, orientation) operators applied as MFMA GEMMs (fmm_upward). */
#define FMM_M2M_NO_TRANSLATION 0
#define FMM_M2M_COMPLEX_TRANSLATION 1
#define FMM_M2M_REAL_TRANSLATION 2
#define FMM_M2M_LAYERS 3
#define FMM_M2M_MATRICES 4

/* stages, in reference order */
int fmm_build_trees(fmm_ctx* ctx);
int fmm_leaf_expansions(fmm_ctx* ctx);
int fmm_upward(fmm_ctx* ctx);
int fmm_upward_with(fmm_ctx* ctx, int m2m_algorithm);   /* calclMultipoleExpansions(M2MAlg, Device) */
int fmm_downward(fmm_ctx* ctx);
int fmm_evaluate_B(fmm_ctx* ctx, double current, double* B);         /* host, m x 3, input order */
int fmm_evaluate_B_dev(fmm_ctx* ctx, double current, double* B);     /* device, m x 3, input order */
/* all stages: build_trees .. evaluate_B_dev */
int fmm_run_B_dev(fmm_ctx* ctx, double current, double* B);
/* reference output order (target-leaf DFS order): order[k] = input index of the k-th result */
int fmm_leaf_order(fmm_ctx* ctx, int64_t* order);

/* vector potential A = current / (4 pi) sum w J / |x - y| (the reference's scaling: no MU0), m x 3,
 * input order; far part from the local expansions, close lists by the near_field rule */
int fmm_evaluate_A(fmm_ctx* ctx, double current, double* A);         /* host */
int fmm_evaluate_A_dev(fmm_ctx* ctx, double current, double* A);     /* device */

/* FastMultipoleSolver::calcBEM (fast_multipole_solver.hpp:47, fast_multipole_solver.cpp:233-250) of BEM
 * sources: sum w ((y - x)(B . n) + (y - x) x (B x n)) / |y - x|^3 / (4 pi) at every target (m x 3, input
 * order; the reference's current argument does not enter) */
int fmm_evaluate_bem(fmm_ctx* ctx, double* out);       /* host */
int fmm_evaluate_bem_dev(fmm_ctx* ctx, double* out);   /* device */

/* treecode at m points (n x 3): B = current MU0 / (4 pi) curl A and A = current / (4 pi) sum M I
 * by M2P walks from the source root (4 R^2 < d^2 accepts a box, failing leaves add nothing), after
 * fmm_upward; B or A may be NULL.  Host and device-pointer variants. */
int fmm_treecode(fmm_ctx* ctx, double current, size_t m, const double* xyz, double* B, double* A);
int fmm_treecode_dev(fmm_ctx* ctx, double current, size_t m, const double* xyz, double* B, double* A);

/* transposed FMM: dLdw[j] = sum_i dLdB[i] . dB_i/dw_j (device pointers, input order) */
int fmm_adjoint_weights_dev(fmm_ctx* ctx, double current, const double* dLdB, double* dLdw);

/* FP64 direct sums over the context's sources at m device targets (B: Biot-Savart, A: vector
 * potential); the sources are split into ranges whose partial sums (a library workspace) are added
 * in range order */
int fmm_direct_B_dev(fmm_ctx* ctx, double current, size_t m, const double* xyz, double* B);
int fmm_direct_A_dev(fmm_ctx* ctx, double current, size_t m, const double* xyz, double* A);

/* fmm_direct_B_dev that allocates nothing (no reference counterpart: for callers that cannot
 * allocate device memory, e.g. code under HIP graph capture or tight serving loops).  The partial
 * sums use a fixed device scratch array of FMM_NOALLOC_SCRATCH_DOUBLES doubles (32 MiB) per GPU,
 * held until the process ends; m <= FMM_NOALLOC_MAX_TARGETS, larger m returns FMM_ERR_LIMIT.
 * Every call on one GPU shares that array: calls that overlap on one GPU (other threads or
 * contexts, other streams) race and must be serialised by the caller. */
#define FMM_NOALLOC_MAX_TARGETS 65536
#define FMM_NOALLOC_SCRATCH_DOUBLES ((size_t)4 << 20)
int fmm_direct_B_noalloc_dev(fmm_ctx* ctx, double current, size_t m, const double* xyz, double* B);

int fmm_get_stats(fmm_ctx* ctx, fmm_stats* stats);

/* ---- parity export (host arrays) ---- */
/* which: 0 source tree, 1 target tree; arrays in canonical BFS node order.  The node arrays are
 * copied from the device on the first export after a build (the device is synchronised). */
int fmm_tree_size(fmm_ctx* ctx, int which, int* num_nodes, int* num_slots);
int fmm_tree_export(fmm_ctx* ctx, int which, int* level, int* parent, int* first_child, int* begin, int* end,
                    int* npts, double* center, int* slots);
int fmm_lists_size(fmm_ctx* ctx, long long* nfar, long long* nclose, long long* nclasses);
int fmm_lists_export(fmm_ctx* ctx, int* far_t, int* far_s, int* close_t, int* close_s);
/* far / near field of the last evaluation, target-leaf order, scaled to tesla */
int fmm_export_B_parts(fmm_ctx* ctx, double* Bfar, double* Bnear);
/* expansion of one node in the reference (unscaled) normalisation: which 0 = multipole of a
 * source node, 1 = local expansion of a target node; out = 3 x (p+1)^2 */
int fmm_export_expansion(fmm_ctx* ctx, int which, int node, double* out);
/* one real-basis translation operator T (row = output), built by the GPU table generator:
 * kind 0 M2M, 1 M2L, 2 L2L; out = (p+1)^2 x (p+1)^2 row-major */
int fmm_make_operator(int kind, int order, const double* t, double sig_in, double sig_out, int m2l_full,
                      double* out);

#ifdef __cplusplus
}
#endif
#endif
