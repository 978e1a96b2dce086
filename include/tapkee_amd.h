/* C-ABI of the MI355X (gfx950) t-SNE engine, libtapkee_amd.so.
 *
 * Plain C types only: row-major arrays, int32 indices, explicit sizes.  Every
 * entry point returns a tka_status; tka_last_error() gives the message of the
 * last failure on the calling thread.  Host-pointer entry points copy in and
 * out and synchronise before returning; *_dev entry points take device
 * pointers and a hipStream_t (as void*) and are asynchronous on that stream
 * except where a scalar result is returned through a host pointer.
 *
 * Reference: tapkee 1.3.4, include/tapkee/external/barnes_hut_sne/tsne.hpp
 * (cited below as tsne.hpp), quadtree.hpp, vptree.hpp.
 */
#ifndef TAPKEE_AMD_H
#define TAPKEE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define TKA_API __attribute__((visibility("default")))
#else
#define TKA_API
#endif

/* Status codes map 1:1 onto tapkee's exception types
 * (include/tapkee/exceptions.hpp:15-92). */
typedef enum tka_status {
    TKA_OK = 0,
    TKA_E_WRONG_PARAMETER = 1,   /* tapkee::wrong_parameter_error   */
    TKA_E_NO_DATA = 2,           /* tapkee::no_data_error           */
    TKA_E_NOT_ENOUGH_MEMORY = 3, /* tapkee::not_enough_memory_error */
    TKA_E_CANCELLED = 4,         /* tapkee::cancelled_exception     */
    TKA_E_UNSUPPORTED = 5,       /* tapkee::unsupported_method_error */
    TKA_E_DEVICE = 6,            /* HIP runtime failure (no tapkee counterpart) */
    TKA_E_INTERNAL = 7
} tka_status;

/* Arithmetic of the kNN distance tiles (the re-rank is always FP64). */
typedef enum tka_dist_dtype {
    TKA_DIST_FP32 = 0, /* FP32 tiles (v_mfma_f32_32x32x2_f32), FP32 accumulate */
    TKA_DIST_BF16 = 1, /* BF16 MFMA tiles, FP32 accumulate */
    TKA_DIST_FP8 = 2   /* FP8 e4m3 MFMA tiles, per-row absmax scaling; approximate:
                          rows failing the error bound are counted, not re-solved */
} tka_dist_dtype;

typedef enum tka_precision { TKA_FP32 = 0, TKA_FP64 = 1 } tka_precision;

typedef enum tka_y0_mode {
    TKA_Y0_RAND_COMPAT = 0, /* std::rand() global state, exactly as tapkee
                               (tsne.hpp:137-138 after the VP-tree's draws,
                               vptree.hpp:226) */
    TKA_Y0_SEEDED = 1,      /* srand(seed) first, then as above */
    TKA_Y0_EXPLICIT = 2     /* caller supplies y0 (N x no_dims) */
} tka_y0_mode;

typedef enum tka_exact_p_mode {
    TKA_EXACT_P_DENSE = 0, /* theta == 0: dense P over all N-1 points (tsne.hpp:92-109) */
    TKA_EXACT_P_KNN = 1    /* theta == 0: kNN P (K = 3*perplexity) + exact repulsion */
} tka_exact_p_mode;

typedef enum tka_dd_mode {
    TKA_DD_TEXTBOOK = 0,      /* ||x-y||^2 */
    TKA_DD_TAPKEE_COMPAT = 1  /* -2<x,y>, the overwrite at tsne.hpp:948-950 */
} tka_dd_mode;

typedef enum tka_kl_mode {
    TKA_KL_AS_REFERENCE = 0, /* BH estimate of Z at theta>0 (tsne.hpp:454-486),
                                dense formula at theta==0 (:409-452) */
    TKA_KL_EXACT_Z = 1       /* exact Z over all pairs, FP64 */
} tka_kl_mode;

/* Options of tka_tsne_run.  tka_default_options() fills the reference's
 * hard-coded schedule (tsne.hpp:71-73, :129, :116). */
typedef struct tka_tsne_options {
    int32_t struct_size;      /* sizeof(tka_tsne_options) */
    int32_t max_iter;         /* 1000 */
    int32_t stop_lying_iter;  /* 250  */
    int32_t mom_switch_iter;  /* 250  */
    double momentum;          /* 0.5  */
    double final_momentum;    /* 0.8  */
    double eta;               /* 200  */
    double exaggeration;      /* 12   */
    int32_t knn_k;            /* 0 -> (int)(3 * perplexity) */
    int32_t dist_dtype;       /* tka_dist_dtype, default BF16 */
    int32_t candidates;       /* k' candidates kept per row before re-rank; 0 -> 1.5 K (FP32 tiles), 5/3 K
                                 (BF16 tiles centred per query block, D <= 158), 1.5 K (BF16 beyond), 2 K (FP8) */
    int32_t loop_precision;   /* tka_precision of the gradient loop, default FP32 */
    int32_t exact_p_mode;     /* tka_exact_p_mode, default DENSE (as the reference) */
    int32_t dd_mode;          /* tka_dd_mode, default TEXTBOOK */
    int32_t y0_mode;          /* tka_y0_mode, default RAND_COMPAT */
    uint32_t seed;            /* for TKA_Y0_SEEDED */
    const double* y0;         /* for TKA_Y0_EXPLICIT, N x no_dims */
    int32_t log_every;        /* KL evaluation period, 50; 0 = never */
    int32_t kl_mode;          /* tka_kl_mode for the periodic evaluations */
    int32_t verbose;          /* print the reference's log lines to stderr */
    int32_t device;           /* HIP device ordinal */
    int64_t device_mem_cap;   /* bytes; 0 = unlimited (forces NOT_ENOUGH_MEMORY) */
    int32_t (*cancel)(void);  /* nonzero -> TKA_E_CANCELLED (checked before work,
                                 as include/tapkee/methods.hpp:30-31) */
} tka_tsne_options;

#define TKA_MAX_KL_TRACE 64

typedef struct tka_tsne_stats {
    double ms_normalize;
    double ms_knn;
    double ms_perplexity;
    double ms_symmetrize;
    double ms_loop;
    double ms_total;
    int64_t nnz;                 /* entries of the symmetric P */
    int32_t knn_k;
    int32_t bisection_unconverged;
    int32_t certificate_failures; /* kNN rows re-solved exactly (see tka_knn) */
    int32_t kl_count;
    int32_t kl_iter[TKA_MAX_KL_TRACE];
    double kl_value[TKA_MAX_KL_TRACE];
    double kl_final_exact;       /* FP64 exact-Z KL of the final embedding */
    double ms_upload;            /* host-to-device copy of X (included in ms_normalize) */
} tka_tsne_stats;

/* Logging, as the reference's Logging singleton and LoggerImplementation
 * (include/tapkee/utils/logging.hpp:45-144): five levels, each enabled or
 * not; a message of an enabled level goes to the installed function, by
 * default one that writes "[level] message" to stderr.  The engine emits
 * the reference's t-SNE messages (tsne.hpp:66,68,191,605) and, at the
 * benchmark level, "Input similarities computation took <s> seconds." and
 * "Main t-SNE loop took <s> seconds." (tsne.hpp:87,142; utils/time.hpp:29-42).
 * A call whose options set `verbose` emits its info and benchmark messages
 * whatever the levels. */
typedef enum tka_log_level {
    TKA_LOG_INFO = 0,
    TKA_LOG_WARNING = 1,
    TKA_LOG_DEBUG = 2,
    TKA_LOG_ERROR = 3,
    TKA_LOG_BENCHMARK = 4
} tka_log_level;
#define TKA_LOG_DEFAULT_LEVELS ((1u << TKA_LOG_WARNING) | (1u << TKA_LOG_ERROR)) /* as logging.hpp:96-98 */
typedef void (*tka_log_fn)(int32_t level, const char* message, void* user);
/* fn == NULL restores the stderr logger.  enabled_levels: bit (1 << level)
 * per enabled level.  Process-wide; fn may be called from any thread that
 * calls the engine. */
TKA_API void tka_set_logger(tka_log_fn fn, void* user, uint32_t enabled_levels);
/* Emits one message through the installed logger if its level is enabled
 * (for front ends that log as the reference's Logging::message_* does). */
TKA_API void tka_log(int32_t level, const char* message);

TKA_API void tka_default_options(tka_tsne_options* opt);
TKA_API const char* tka_last_error(void);
TKA_API int32_t tka_version(void);

/* Full t-SNE.  Replaces tsne::TSNE::run(X, N, D, Y, no_dims, perplexity,
 * theta) (tsne.hpp:61-205) with its validation
 * (include/tapkee/methods/tsne.hpp:18-22, methods/base.hpp:38-43).
 * X: N x D row-major FP64 (tapkee's D x N column-major DenseMatrix), not
 * modified.  Y: N x no_dims row-major output.  opt/stats may be NULL. */
TKA_API tka_status tka_tsne_run(const double* X, int32_t N, int32_t D, double* Y, int32_t no_dims,
                                double perplexity, double theta, const tka_tsne_options* opt,
                                tka_tsne_stats* stats);

/* Sessions: the same computation as tka_tsne_run split into its setup
 * (tsne.hpp:61-139: validation, normalisation, P, Y0) and steps of the main
 * loop (tsne.hpp:143-193), so callers can advance, inspect and time
 * iterations.  One session owns one HIP stream and its device buffers.
 * Steps are asynchronous except at KL evaluations (every log_every
 * iterations), which read a scalar back. */
typedef struct tka_session tka_session;

typedef struct tka_kernel_timing {
    int64_t launches_build;      /* tree builds (keys, sort, nodes) timed; while timing, the build runs
                                    without the attraction rows it otherwise carries */
    int64_t launches_repulsion;  /* Barnes-Hut traversals timed (a timed step runs its traversal 4 times
                                    back to back between one pair of events) */
    int64_t launches_step;       /* optimiser updates timed (with the attraction fused when theta = 0) */
    int64_t launches_exact;      /* exact all-pairs repulsion launches timed */
    double ms_build;             /* summed HIP-event durations on the session stream */
    double ms_repulsion;
    double ms_step;
    double ms_exact;
    int64_t interactions;        /* Barnes-Hut cell/point interactions in the timed traversals */
    int64_t nnz;                 /* entries of P (edges per attraction pass) */
    int32_t n;
    int32_t dims;
    int64_t launches_attraction; /* Barnes-Hut loop: attraction passes timed as launches of their own
                                    (untimed steps run these rows inside the tree build's launches) */
    double ms_attraction;
} tka_kernel_timing;

TKA_API tka_status tka_session_create(const double* X, int32_t N, int32_t D, int32_t no_dims, double perplexity,
                                      double theta, const tka_tsne_options* opt, tka_session** out);
/* Sessions from a precomputed similarity structure instead of features
 * (the counterpart of tapkee's precomputed callbacks,
 * include/tapkee/callbacks/precomputed_callbacks.hpp:20-42, for t-SNE):
 *  - tka_session_create_knn: a neighbour graph, idx (N x K int32, the
 *    object itself excluded) and dist (N x K squared distances); the
 *    perplexity calibration, symmetrisation and division by the sum follow
 *    (tsne.hpp:601-718, :207-322, :119-124);
 *  - tka_session_create_p: a joint P in CSR (row_ptr N+1, col and val
 *    row_ptr[N] entries), used after division by its sum (tsne.hpp:119-124);
 *    it should be symmetric, as the reference's is.
 * Both give a sparse P also at theta == 0 (exact repulsion, CSR attraction).
 * Indices, distances and values are checked on the host (out of range, the
 * object itself or a neighbour listed twice in a row, negative or non-finite
 * -> TKA_E_WRONG_PARAMETER).  Y0 as in the options;
 * TKA_Y0_RAND_COMPAT draws from std::rand() with no VP-tree draws before. */
TKA_API tka_status tka_session_create_knn(const int32_t* idx, const double* dist, int32_t N, int32_t K, int32_t no_dims,
                                          double perplexity, double theta, const tka_tsne_options* opt,
                                          tka_session** out);
TKA_API tka_status tka_session_create_p(const int32_t* row_ptr, const int32_t* col, const double* val, int32_t N,
                                        int32_t no_dims, double theta, const tka_tsne_options* opt, tka_session** out);
TKA_API tka_status tka_session_step(tka_session* s, int32_t iterations);
TKA_API tka_status tka_session_synchronize(tka_session* s);
/* Current embedding, zero-mean, N x no_dims FP64. */
TKA_API tka_status tka_session_embedding(tka_session* s, double* Y);
/* KL(P||Q) of the current embedding with the exact normaliser (FP64). */
TKA_API tka_status tka_session_kl(tka_session* s, double* kl);
TKA_API tka_status tka_session_stats(tka_session* s, tka_tsne_stats* stats);
/* enable != 0: record HIP events around every kernel class in later steps. */
TKA_API tka_status tka_session_profile(tka_session* s, int32_t enable);
/* Test hook for the error path of the Barnes-Hut tree build: marks the
 * session's Morton sort as failed, as a tile look-back that gave up would
 * (the sort itself stays valid).  Later steps skip the repulsion walk, and
 * tka_session_embedding / tka_session_kl return TKA_E_INTERNAL.  No effect
 * at theta == 0.  (No counterpart in the reference, whose tree has no
 * device sort.) */
TKA_API tka_status tka_session_inject_sort_fault(tka_session* s);
/* Synchronises, then returns and clears the accumulated timings. */
TKA_API tka_status tka_session_kernel_timing(tka_session* s, tka_kernel_timing* t);
TKA_API void tka_session_destroy(tka_session* s);

/* Normalisation: zeroMean + division by the signed maximum coefficient
 * (tsne.hpp:88-89, :488-518).  In place on a host N x D array. */
TKA_API tka_status tka_normalize(double* X, int32_t N, int32_t D);

/* Brute-force K nearest neighbours on squared Euclidean distance, self
 * excluded, rows sorted by (FP64 distance, index).  Replaces the VP-tree
 * search (tsne.hpp:628-646, vptree.hpp:126-293) with an exact answer.
 * candidates = k' kept from the low-precision tiles (0 -> the default of
 * tka_tsne_options.candidates).
 * certificate_failures (may be NULL) counts rows whose error-bound test
 * failed and were re-solved in FP64. */
TKA_API tka_status tka_knn(const double* X, int32_t N, int32_t D, int32_t K, int32_t dist_dtype,
                           int32_t candidates, int32_t* idx, double* dist, int32_t* certificate_failures);

/* Connectivity of a neighbour graph as tapkee's is_connected
 * (include/tapkee/neighbors/connected.hpp:19-61): depth-first search from
 * object 0 along the directed lists idx (N x k); *connected = 1 when every
 * object is reached.  Host only.  With tka_knn it makes tapkee's
 * find_neighbors (neighbors.hpp:172-216) for feature vectors under the
 * Euclidean distance: exact neighbours replace the Brute, VpTree and
 * CoverTree searches (include/tapkee_amd/neighbors.hpp). */
TKA_API tka_status tka_neighbors_connected(const int32_t* idx, int32_t N, int32_t k, int32_t* connected);

/* Perplexity calibration of kNN rows: bisection on beta in FP64
 * (tsne.hpp:648-711).  dist: N x K squared distances.  val: N x K row
 * normalised P.  beta/iters may be NULL.  Returns through *unconverged the
 * number of rows that hit the 200-evaluation cap (may be NULL). */
TKA_API tka_status tka_perplexity(const double* dist, int32_t N, int32_t K, double perplexity, double* val,
                                  double* beta, int32_t* iters, int32_t* unconverged);

/* Dense calibration over all N-1 points (tsne.hpp:520-599) with the
 * distance convention dd_mode.  X: N x D (already normalised).  P: N x N. */
TKA_API tka_status tka_perplexity_dense(const double* X, int32_t N, int32_t D, double perplexity,
                                        int32_t dd_mode, double* P);

/* Symmetrisation (tsne.hpp:207-322): P_sym = (P + P^T)/2 over the union
 * pattern, CSR with columns ascending within each row.  idx/val: N x K.
 * row_ptr: N+1; col/sval: capacity 2*N*K.  *nnz receives the entry count. */
TKA_API tka_status tka_symmetrize(const int32_t* idx, const double* val, int32_t N, int32_t K,
                                  int32_t* row_ptr, int32_t* col, double* sval, int64_t* nnz);

/* The whole input-similarity stage of the Barnes-Hut path: normalise X,
 * exact kNN (K = knn_k, 0 -> (int)(3 * perplexity)), perplexity calibration,
 * symmetrisation and division by the sum (tsne.hpp:86-125, without the x12
 * of :131-134).  X: N x D (not modified).  row_ptr: N+1; col/val: capacity
 * 2*N*K; *nnz receives the entry count. */
TKA_API tka_status tka_build_p(const double* X, int32_t N, int32_t D, double perplexity, int32_t knn_k,
                               int32_t dist_dtype, int32_t* row_ptr, int32_t* col, double* val, int64_t* nnz);

/* t-SNE gradient at a fixed Y from a CSR P, without the factor 4, as
 * tsne.hpp:325-351 (theta > 0, Barnes-Hut, 2-D only) or the exact sum over
 * all pairs (theta == 0).  precision: tka_precision.  *sum_q receives Z. */
TKA_API tka_status tka_gradient(const double* Y, int32_t N, int32_t no_dims, const int32_t* row_ptr,
                                const int32_t* col, const double* val, double theta, int32_t precision,
                                double* dY, double* sum_q);

/* Exact gradient with a dense P (tsne.hpp:353-407), dd_mode as above. */
TKA_API tka_status tka_gradient_dense(const double* P, const double* Y, int32_t N, int32_t no_dims,
                                      int32_t dd_mode, int32_t precision, double* dY, double* sum_q);

/* KL(P||Q).  Sparse P: mode TKA_KL_AS_REFERENCE is tsne.hpp:454-486
 * (theta > 0), TKA_KL_EXACT_Z the exact normaliser.  Always FP64. */
TKA_API tka_status tka_kl(const double* Y, int32_t N, int32_t no_dims, const int32_t* row_ptr,
                          const int32_t* col, const double* val, double theta, int32_t kl_mode, double* kl);

/* Dense-P KL (tsne.hpp:409-452), dd_mode as above. */
TKA_API tka_status tka_kl_dense(const double* P, const double* Y, int32_t N, int32_t no_dims, int32_t dd_mode,
                                double* kl);

/* One optimiser step: gains, momentum, update, zeroMean (tsne.hpp:152-167). */
TKA_API tka_status tka_update(const double* dY, double* uY, double* gains, double* Y, int32_t N,
                              int32_t no_dims, double momentum, double eta, int32_t precision);

/* Barnes-Hut tree export (pre-order, compressed): node count in *n_nodes;
 * per node level, count, leaf point (-1 for internal), skip index and centre
 * of mass.  Arrays need capacity 2*N.  root_box: centre x, y, half-widths
 * hw, hh (quadtree.hpp:99-134).  The tree is built twice, the second time
 * with the radix-sort width the first build's key statistics select, and
 * TKA_E_INTERNAL is returned unless both agree node for node. */
TKA_API tka_status tka_bh_tree(const double* Y, int32_t N, int32_t precision, int32_t* n_nodes, int32_t* level,
                               int32_t* count, int32_t* leaf_point, int32_t* skip, double* com_x, double* com_y,
                               double* root_box);

/* Device-pointer variants for PyTorch (autograd) -------------------------- */

/* FP32 gradient on device arrays: Y (N x 2 or N x no_dims), CSR P in FP32
 * values.  dY (device) receives pos - neg/Z (no factor 4).  *sum_q (host)
 * receives Z; the call synchronises the stream to return it. */
TKA_API tka_status tka_gradient_dev(const float* Y, int32_t N, int32_t no_dims, const int32_t* row_ptr,
                                    const int32_t* col, const float* val, double theta, float* dY, double* sum_q,
                                    void* stream);

/* FP32 sparse KL on device arrays with an exact (theta == 0) or BH Z. */
TKA_API tka_status tka_kl_dev(const float* Y, int32_t N, int32_t no_dims, const int32_t* row_ptr,
                              const int32_t* col, const float* val, double theta, double* kl, double* sum_q,
                              void* stream);

/* KL(P||Q) and its gradient in one call on device arrays: the forward of
 * the autograd function, which keeps the gradient for its backward.  theta
 * == 0: one all-pairs pass gives the repulsion and Z, and the KL
 * (sum p log(p Z (1 + d^2)) over P's pattern) uses that Z; theta > 0: the
 * Barnes-Hut tree gives both.  dY receives pos - neg/Z (no factor 4, as
 * tsne.hpp:325-407).  precision TKA_FP32: Y, val, dY are float; TKA_FP64:
 * double.  *kl and *sum_q (may be NULL) on the host; synchronises the
 * stream. */
TKA_API tka_status tka_kl_grad_dev(const void* Y, int32_t N, int32_t no_dims, const int32_t* row_ptr,
                                   const int32_t* col, const void* val, double theta, int32_t precision, double* kl,
                                   void* dY, double* sum_q, void* stream);

/* Allocation-free variant of tka_gradient_dev: the same arguments, kernels
 * and results (bit for bit), with every buffer beyond the inputs and outputs
 * taken from a fixed device region that the library holds on each GPU it
 * runs on (about 20 MB, until the process ends) instead of allocated.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).
 * N <= TKA_NOALLOC_MAX_POINTS, else TKA_E_WRONG_PARAMETER.  The region's
 * address is looked up on the first call on a device (that call is the
 * set-up; later calls allocate nothing).  Calls of one process are
 * serialised and each synchronises its stream; the region keeps the last
 * call's contents. */
#define TKA_NOALLOC_MAX_POINTS 131072
TKA_API tka_status tka_gradient_dev_noalloc(const float* Y, int32_t N, int32_t no_dims, const int32_t* row_ptr,
                                            const int32_t* col, const float* val, double theta, float* dY,
                                            double* sum_q, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* TAPKEE_AMD_H */
