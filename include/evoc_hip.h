/* evoc_hip — MI355X (gfx950) implementation of evoc-rs's clustering pipeline.
 *
 * C ABI: every function is extern "C", never throws, returns an int status
 * (EVOC_OK = 0; otherwise evoc_hip_last_error() holds a thread-local message).
 * Caller buffers are borrowed for the duration of a call and never retained.
 * Host-pointer entry points accept ordinary host memory; evoc_hip_run_device takes
 * the input matrix already resident in device memory. Calls are serialised inside
 * the library (one device, one stream per device).
 *
 * The types mirror the reference's public interface:
 *   evoc_params_t          <- EvocParams<T>            (src/lib.rs:52-99)
 *   evoc_hip_nn_params_t   <- NearestNeighbourParamsEvoc.dist_metric
 *                             (src/nearest_neighbours/nearest_neighbour_cpu.rs:43-181)
 *                             plus the device knobs this implementation adds
 *   evoc_result_t          <- EvocResult<T>            (src/lib.rs:106-161)
 */
#ifndef EVOC_HIP_H
#define EVOC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  EVOC_OK = 0,
  EVOC_ERR_INVALID_ARGUMENT = 1,
  EVOC_ERR_HIP = 2,
  EVOC_ERR_NO_DEVICE = 3,
  EVOC_ERR_INTERNAL = 4
};

enum { EVOC_METRIC_EUCLIDEAN = 0, EVOC_METRIC_COSINE = 1 };
/* kNN arithmetic: AUTO picks one bf16 MFMA per product when every input value is
 * bf16-representable, else the three-product hi/lo split; both are followed by the
 * exact f32 re-rank, so the returned neighbours are identical in every mode. FP8
 * (cosine metric only) runs the first pass on FP8-e4m3 values with per-row scales
 * (max|x| / 448) and f32 accumulation; its certificate carries the quantisation error
 * of the worst row, and rows it cannot certify are searched again in bf16 (then
 * exactly), so the neighbours are again identical. */
enum { EVOC_KNN_AUTO = 0, EVOC_KNN_SPLIT = 1, EVOC_KNN_BF16 = 2, EVOC_KNN_FP8 = 3 };

/* Mirrors EvocParams (src/lib.rs:52-80); defaults src/lib.rs:83-99. */
typedef struct {
  int32_t n_neighbours;            /* 15 */
  float noise_level;               /* 0.5 */
  int32_t n_epochs;                /* 50 */
  int32_t embedding_dim;           /* 0 = None -> clamp(n_neighbours / 4, 4, 16) */
  float neighbour_scale;           /* 1.0 */
  int32_t symmetrise;              /* 1 */
  int32_t min_samples;             /* 5 */
  int32_t base_min_cluster_size;   /* 5 */
  int32_t approx_n_clusters;       /* 0 = None */
  double min_similarity_threshold; /* 0.2 */
  int32_t max_layers;              /* 10 */
} evoc_params_t;

/* Nearest-neighbour and device settings (dist_metric of NearestNeighbourParamsEvoc,
 * src/nearest_neighbours/nearest_neighbour_cpu.rs:45, plus this library's knobs). */
typedef struct {
  int32_t metric;          /* EVOC_METRIC_* */
  int32_t knn_precision;   /* EVOC_KNN_* */
  int32_t warm_start;      /* 0 = label-propagation init (plain), 1 = learned MLP init,
                            * accepted only if it passes the check on its training
                            * subsample, else the plain init runs */
  int32_t warm_epochs;     /* embedding epochs of the warm-started run; 0 = 30% of n_epochs */
  int32_t mlp_subsample;   /* points in the MLP training subsample; 0 = 50000 */
  int32_t device_id;       /* HIP device ordinal */
} evoc_hip_nn_params_t;

/* Mirrors EvocResult (src/lib.rs:106-118). Owned by the library; release with
 * evoc_hip_result_free. Layers are sorted finest (most clusters) first. */
typedef struct {
  int64_t n;
  int32_t k;
  int32_t n_layers;
  int32_t dim;
  int32_t best_layer;              /* EvocResult::best_labels index (src/lib.rs:124-137) */
  int64_t* cluster_layers;         /* n_layers * n, -1 = noise */
  float* membership_strengths;     /* n_layers * n */
  double* persistence_scores;      /* n_layers */
  int32_t* nn_indices;             /* n * k, excluding self */
  float* nn_distances;             /* n * k */
  float* embedding;                /* n * dim, the node embedding clustered */
  double stage_ms[8];              /* knn, fuzzy, init, embed, mst, layers, total, upload */
  int64_t knn_fallback_rows;       /* rows re-searched exactly after a failed certificate */
  double knn_kernel_ms;            /* MFMA top-candidate kernel, hipEvent-timed on its stream */
  double embed_kernel_ms;          /* all embedding epochs, hipEvent-timed */
  /* learned warm start (warm_start = 1): whether its layout was accepted, and the check
   * on the training subsample that decided it (warm-started short run against the plain
   * run: best-layer ARI >= 0.98 and MST total weight ratio within 1%) */
  int32_t warm_start_used;
  double warm_start_ari;
  double warm_start_mst_ratio;
} evoc_result_t;

void evoc_hip_default_params(evoc_params_t* p);
void evoc_hip_default_nn_params(evoc_hip_nn_params_t* p);

/* evoc() (src/lib.rs:199-352) on a host matrix with arbitrary strides (elements).
 * pre_idx / pre_dist: PreComputedKnn (src/prelude.rs:85), n * pre_k each, or NULL. */
int evoc_hip_run(const float* data, int64_t n, int64_t d, int64_t row_stride,
                 int64_t col_stride, const int32_t* pre_idx, const float* pre_dist,
                 int32_t pre_k, const evoc_params_t* params, const evoc_hip_nn_params_t* nn,
                 uint64_t seed, int32_t verbose, evoc_result_t** out);
/* Same, with data a row-major n x d device pointer (leading dimension ld). */
int evoc_hip_run_device(const float* d_data, int64_t n, int64_t d, int64_t ld,
                        const evoc_params_t* params, const evoc_hip_nn_params_t* nn,
                        uint64_t seed, int32_t verbose, evoc_result_t** out);
void evoc_hip_result_free(evoc_result_t* r);
const char* evoc_hip_last_error(void);
int evoc_hip_device_count(int32_t* count);
/* Frees cached device memory held between calls. */
int evoc_hip_release(void);

/* ---- stage entry points (host pointers in and out), each one reference function ---- */

/* Exhaustive arm of run_ann_search (src/nearest_neighbours/nearest_neighbour_cpu.rs:
 * 225-364) with the j != i self filter (nearest_neighbour_gpu.rs:291-303): top-k per
 * row by (distance, index). Euclidean distance = sqrt of squared L2; cosine = 1-cos. */
int evoc_hip_knn(const float* x, int64_t n, int64_t d, int32_t metric, int32_t k,
                 int32_t precision, int32_t* idx_out, float* dist_out, int64_t* fallback_rows,
                 double* kernel_ms);

/* Diagnostics of the calling thread's last kNN search (evoc_hip_knn or the kNN stage of
 * evoc_hip_run / evoc_hip_run_device): the MFMA flop the tile kernel executed, its
 * hipEvent time in ms, whether the symmetric sweep (each pair of 256-row blocks once,
 * transposed keys queued for the other block) ran, and how many rows whose certificate
 * failed were searched again with 64 candidates (escalated) before the exact scan took
 * the rest. No counterpart in the reference: it serves roofline accounting and tests.
 * Any pointer may be NULL. */
int evoc_hip_knn_last_stats(double* kernel_flop, double* kernel_ms, int32_t* symmetric,
                            int64_t* escalated_rows);

/* The first tile pass of evoc_hip_knn alone, for tests of the tile kernels: per row the
 * raw candidate lists (*ncand entries: two column halves, columns c with (c % 128) < 64
 * and >= 64 of each 128-column tile, each ascending by the pass's key), the keys, the
 * per-row key scale (cosine: 1/||x||, or s/||x|| with the FP8 row scale s = max|x|/448;
 * the pass's key of a pair is -(dot * scale of the column), the dot of the stored values)
 * and the certificate's relative error term of the pass. cand_idx and cand_key need room
 * for n * 32 entries. No counterpart in the reference. Supported for the streaming
 * kernel's precisions: EVOC_KNN_SPLIT and EVOC_KNN_FP8. */
int evoc_hip_knn_tile_pass(const float* x, int64_t n, int64_t d, int32_t metric, int32_t k,
                           int32_t precision, int32_t* cand_idx, float* cand_key,
                           float* key_scale, int32_t* ncand, float* gamma);

/* build_fuzzy_simplicial_set (src/graph/fuzzy_graph.rs:275-295) as CSR. indices and
 * weights need room for 2*n*k entries; *nnz receives the stored count. */
int evoc_hip_fuzzy(const int32_t* knn_idx, const float* knn_dist, int64_t n, int32_t k,
                   float effective_k, int32_t symmetrise, float* sigma, float* rho,
                   int64_t* indptr, int32_t* indices, float* weights, int64_t* nnz);

/* evoc_embedding (src/graph/embedding.rs:48-298) on a CSR graph. init may be NULL
 * (random U(-0.25, 0.25) layout). */
int evoc_hip_embed(const int64_t* indptr, const int32_t* indices, const float* weights,
                   int64_t n, int32_t dim, int32_t n_epochs, float noise_level,
                   float negative_sample_rate, float initial_alpha, float beta1, float beta2,
                   float eps, const float* init, uint64_t seed, float* out);

/* Allocation-free variant of evoc_hip_embed, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops). No counterpart in the
 * reference. Every pointer is device memory: the graph must be symmetric with each row
 * sorted by column (the fuzzy set's CSR; its rows are then the embedding's slot lists),
 * with finite weights; d_emb holds the initial layout (or is filled from the seed when
 * random_init != 0) and receives the final one, n x dim row-major. The epochs' scratch is
 * a static per-GPU device array: n <= 65536, nnz <= 2097152, dim <= 16 (larger sizes
 * return EVOC_ERR_INVALID_ARGUMENT). Runs on the library's stream for device_id and
 * returns after it finishes. Calls that overlap on one GPU share the scratch and race;
 * the library's call lock serialises calls made through it. */
int evoc_hip_embed_noalloc(const int64_t* d_indptr, const int32_t* d_indices,
                           const float* d_weights, int64_t n, int64_t nnz, int32_t dim,
                           int32_t n_epochs, float noise_level, float negative_sample_rate,
                           float initial_alpha, float beta1, float beta2, float eps,
                           int32_t random_init, uint64_t seed, int32_t device_id, float* d_emb);

/* label_propagation_init (src/graph/label_prop.rs:537-564) on a column-sorted CSR.
 * data (n x d, row-major) may be NULL. */
int evoc_hip_label_prop_init(const int64_t* indptr, const int32_t* indices,
                             const float* weights, int64_t n, int32_t dim, const float* data,
                             int64_t d, uint64_t seed, float* out);

/* Squared core distances (src/clustering/mst.rs:92-105). */
int evoc_hip_core_sq(const float* emb, int64_t n, int32_t dim, int32_t min_samples,
                     float* core_sq);

/* build_mst (src/clustering/mst.rs:82-166): n-1 edges in (w, u, v) order, u < v,
 * weights = sqrt of the squared mutual-reachability distance. */
int evoc_hip_mst(const float* emb, int64_t n, int32_t dim, int32_t min_samples, int32_t* u,
                 int32_t* v, float* w);

/* condense_tree (src/clustering/condensed_tree.rs:127-248) of n - 1 single-linkage rows
 * (src/clustering/linkage.rs:301-310: row i is node n + i, left/right child ids, distance,
 * size), on the device. The rows come back in the reference's order into parent, child,
 * lambda and child_size, which need 2n entries; *n_rows receives their number. */
int evoc_hip_condense_tree(const int64_t* left, const int64_t* right, const float* distance,
                           const int64_t* size, int64_t n, int64_t min_cluster_size,
                           int64_t* parent, int64_t* child, float* lambda, int64_t* child_size,
                           int64_t* n_rows);

/* build_cluster_layers (src/clustering/persistence.rs:387-463). labels/strengths
 * need max_layers * n entries, persistence max_layers; *n_layers receives the count. */
int evoc_hip_cluster_layers(const float* emb, int64_t n, int32_t dim, int32_t min_samples,
                            int64_t base_min_cluster_size, double min_similarity_threshold,
                            int64_t max_layers, int64_t* labels, float* strengths,
                            double* persistence, int64_t* n_layers);

/* search_for_n_clusters (src/lib.rs:377-447). */
int evoc_hip_search_n_clusters(const float* emb, int64_t n, int32_t dim, int32_t min_samples,
                               int64_t target_k, int64_t* labels, float* strengths);

/* Learned warm start (no counterpart in the reference): one forward/backward pass of
 * the MLP  d -> 256 -> 256 -> out  (GELU; bf16 MFMA, fp32 accumulation) for the mean
 * squared error against t (rows x out). Weights are fp32 host arrays, layer l has
 * w[l] (out_l x in_l, row-major) and b[l]; gradients are written in the same layout,
 * y receives the prediction, *loss the MSE. Used to check the backward kernels. */
int evoc_hip_mlp_loss_grads(const float* x, const float* t, int32_t rows, int32_t d,
                            int32_t out, const float* w0, const float* b0, const float* w1,
                            const float* b1, const float* w2, const float* b2, float* gw0,
                            float* gb0, float* gw1, float* gb1, float* gw2, float* gb2,
                            float* y, double* loss);

#ifdef __cplusplus
}
#endif

#endif /* EVOC_HIP_H */
