// evoc oracle: plain C++17 (+OpenMP) CPU restatement of evoc-rs 0.2.7's clustering
// pipeline. TEST INFRASTRUCTURE ONLY: it is the checker that the HIP library is
// compared against, and the CPU timing column. The product library never links,
// loads or calls it.
//
// Every function cites the reference file:line it restates. Where the reference's
// output depends on an unpinned third-party detail (rand's SmallRng stream,
// FxHashMap iteration order, the `wide` SIMD sum order inside ann-search-rs,
// faer's SVD sign convention) the oracle fixes one canonical choice; the HIP
// library adopts the same one:
//   * f32 distances: 64-way strided partial sums + halving tree (high-dim), plain
//     sequential sum (embedding space);
//   * kNN order (dist, idx); self excluded by j != i;
//   * MST edge order (w^2, min(u,v), max(u,v)); linkage sort (w, min, max);
//   * symmetrised rows sorted by column;
//   * Philox4x32-10 counter-based random numbers keyed by (seed, stream, ...).
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// metric: 0 = euclidean (sqrt of squared L2), 1 = cosine (1 - cos)
int orc_knn_exhaustive(const float* x, int64_t n, int64_t d, int metric, int k,
                       int32_t* idx_out, float* dist_out);
int orc_knn_rows(const float* x, int64_t n, int64_t d, int metric, int k, const int64_t* rows,
                 int64_t nrows, int32_t* idx_out, float* dist_out);
// canonical f32 distance between two rows (same routine the kNN uses)
float orc_canon_dist(const float* a, const float* b, int64_t d, int metric);

// fuzzy simplicial set; returns nnz (graph rows sorted by col when symmetrised).
// indptr has n+1 entries; indices/weights need capacity 2*n*k.
int64_t orc_fuzzy_graph(const int32_t* knn_idx, const float* knn_dist, int64_t n, int k,
                        float effective_k, int symmetrise,
                        float* sigma_out, float* rho_out,
                        int64_t* indptr, int32_t* indices, float* weights);

int orc_embedding(const int64_t* indptr, const int32_t* indices, const float* weights,
                  int64_t n, int dim, int n_epochs, float noise_level,
                  float negative_sample_rate, float initial_alpha, float beta1,
                  float beta2, float eps, const float* init, uint64_t seed, float* out);

int orc_label_prop_init(const int64_t* indptr, const int32_t* indices,
                        const float* weights, int64_t n, int dim, const float* data,
                        int64_t d, uint64_t seed, float* out);

int orc_core_sq(const float* emb, int64_t n, int dim, int min_samples, float* core_sq);
int orc_core_sq_f64(const double* emb, int64_t n, int dim, int min_samples, double* core_sq);

// n-1 edges in canonical (w^2, min, max) order; weights are sqrt'd
int64_t orc_mst(const float* emb, int64_t n, int dim, int min_samples,
                int32_t* u, int32_t* v, float* w);
int64_t orc_mst_f64(const double* emb, int64_t n, int dim, int min_samples,
                    int32_t* u, int32_t* v, double* w);

// linkage rows (left, right, distance, size); input edges are sorted in place
// by the canonical (w, min, max) order
int orc_linkage(int32_t* u, int32_t* v, float* w, int64_t n_edges, int64_t n,
                int64_t* left, int64_t* right, float* dist, int64_t* size);
int orc_linkage_f64(int32_t* u, int32_t* v, double* w, int64_t n_edges, int64_t n,
                    int64_t* left, int64_t* right, double* dist, int64_t* size);

// condensed tree; returns node count (capacity 2n)
int64_t orc_condense(const int64_t* left, const int64_t* right, const float* dist,
                     const int64_t* size, int64_t n_rows, int64_t n, int64_t mcs,
                     int64_t* parent, int64_t* child, float* lambda_val,
                     int64_t* child_size);
int64_t orc_condense_f64(const int64_t* left, const int64_t* right, const double* dist,
                         const int64_t* size, int64_t n_rows, int64_t n, int64_t mcs,
                         int64_t* parent, int64_t* child, double* lambda_val,
                         int64_t* child_size);

int64_t orc_extract_leaves(const int64_t* parent, const int64_t* child,
                           const int64_t* child_size, int64_t n_nodes, int64_t n,
                           int64_t* leaves_out);
int orc_label_vector(const int64_t* parent, const int64_t* child, int64_t n_nodes,
                     const int64_t* clusters, int64_t n_clusters, int64_t n,
                     int64_t* labels);
int orc_strengths(const int64_t* parent, const int64_t* child, const float* lambda_val,
                  const int64_t* child_size, int64_t n_nodes, const int64_t* clusters,
                  int64_t n_clusters, const int64_t* labels, int64_t n, float* out);

int64_t orc_find_peaks(const double* x, int64_t len, int64_t* peaks);

// persistence curve for a condensed tree: fp64 = 1 uses double throughout
// (the HIP library's precision), 0 = f32 as in the reference. Returns
// (number of barcodes << 32) | number of size bins; the bins go to sizes_out and
// persistence_out, the barcodes to birth_out, death_out and lambda_death_out.
int64_t orc_persistence(const int64_t* parent, const int64_t* child, const float* lambda_val,
                        const int64_t* child_size, int64_t n_nodes, int64_t n,
                        int64_t base_mcs, int fp64, double* sizes_out,
                        double* persistence_out, double* birth_out, double* death_out,
                        double* lambda_death_out);

int orc_select_diverse_peaks(const int64_t* peaks, int64_t n_peaks,
                             const double* persistence, const double* sizes,
                             const double* births, const double* deaths, int64_t n_bc,
                             double min_sim, int64_t max_layers, int64_t* selected,
                             int64_t* n_selected);

// layered clustering of an embedding; returns number of layers.
// labels: max_layers*n, strengths: max_layers*n, persistence: max_layers
int64_t orc_cluster_layers(const float* emb, int64_t n, int dim, int min_samples,
                           int64_t base_mcs, double min_sim, int64_t max_layers, int fp64,
                           int64_t* labels, float* strengths, double* persistence);
int64_t orc_cluster_layers_f64(const double* emb, int64_t n, int dim, int min_samples,
                               int64_t base_mcs, double min_sim, int64_t max_layers,
                               int64_t* labels, double* strengths, double* persistence);
int orc_search_n_clusters(const float* emb, int64_t n, int dim, int min_samples,
                          int64_t target_k, int64_t* labels, float* strengths);
int orc_search_n_clusters_f64(const double* emb, int64_t n, int dim, int min_samples,
                              int64_t target_k, int64_t* labels, double* strengths);

typedef struct {
  int32_t n_neighbours;
  float noise_level;
  int32_t n_epochs;
  int32_t embedding_dim;  // 0 = auto
  float neighbour_scale;
  int32_t symmetrise;
  int32_t min_samples;
  int32_t base_min_cluster_size;
  int32_t approx_n_clusters;  // 0 = none
  double min_similarity_threshold;
  int32_t max_layers;
  int32_t metric;
  int32_t persistence_fp64;
} orc_params_t;

// Full pipeline. pre_idx/pre_dist may be NULL (then exhaustive kNN runs).
// Outputs: labels (max_layers*n), strengths (max_layers*n), persistence (max_layers),
// knn_idx/knn_dist (n*k), embedding (n*dim) may each be NULL. timings[5] (seconds per
// stage: knn, fuzzy, init, embed, cluster) may be NULL. Returns number of layers.
int64_t orc_evoc(const float* x, int64_t n, int64_t d, const int32_t* pre_idx,
                 const float* pre_dist, const orc_params_t* p, uint64_t seed,
                 int64_t* labels, float* strengths, double* persistence,
                 int32_t* knn_idx, float* knn_dist, float* embedding, float* init_out,
                 double* timings);

uint32_t orc_philox_u32(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3);
int orc_num_threads(void);

// ---- pieces exported for the known-answer tests and the CPU baselines
int orc_knn_nndescent(const float* x, int64_t n, int64_t d, int metric, int k, uint64_t seed,
                      int max_iters, float delta, float rho, int32_t* idx_out, float* dist_out,
                      int* iters_done);
int64_t orc_fuzzy_graph_f64(const int32_t* knn_idx, const double* knn_dist, int64_t n, int k,
                            float effective_k, int symmetrise, double* sigma_out,
                            double* rho_out, int64_t* indptr, int32_t* indices,
                            double* weights);
int64_t orc_symmetrise(const int64_t* rows, const int32_t* cols, const float* vals, int64_t nnz,
                       int64_t n, int64_t* indptr, int32_t* indices, float* weights);
// op: 0 transpose, 1 product a*b, 2 element-wise product, 3 L2-normalised columns,
// 4 L1-normalised rows, 5 values clipped to [0, 1] (b unused except for 1 and 2);
// returns nnz; -1 for an unknown op, or when cap is too small (shape then holds the
// result's shape)
int64_t orc_csr_op(int op, const int64_t* aip, const int64_t* aix, const float* adv, int64_t anr,
                   int64_t anc, const int64_t* bip, const int64_t* bix, const float* bdv,
                   int64_t bnr, int64_t bnc, int64_t* oip, int64_t* oix, float* odv,
                   int64_t cap, int64_t* shape);
int64_t orc_csr_from_coo(const int64_t* rows, const int64_t* cols, const float* vals,
                         int64_t nnz, int64_t n, int64_t* oip, int64_t* oix, float* odv);
int orc_csr_from_partition(const int64_t* part, int64_t n, int64_t n_parts, int64_t* oip,
                           int64_t* oix, float* odv);
int orc_csr_matmul_dense(const int64_t* ip, const int64_t* ix, const float* dv, int64_t nr,
                         int64_t nc, const float* rhs, int64_t d, float* out);
int orc_lp_iteration(const int64_t* ip, const int64_t* ix, const float* dv, int64_t n,
                     const int64_t* labels, int64_t* out);
int orc_lp_outliers(const int64_t* ip, const int64_t* ix, const float* dv, int64_t n,
                    int64_t* labels, uint64_t seed);
int64_t orc_lp_remap(int64_t* labels, int64_t n);
int64_t orc_lp_loop(const int64_t* ip, const int64_t* ix, const float* dv, int64_t n,
                    int64_t approx_parts, int n_iter, uint64_t seed, int64_t* part);
int orc_unit_range(float* m, int64_t n, int64_t d);
int orc_pca_init(const float* data, int64_t n, int64_t d, int dim, float* out);
int orc_random_init(int64_t n, int dim, uint64_t seed, float* out);
int orc_lp_expand(const int64_t* ip, const int64_t* ix, const float* dv, int64_t n,
                  const int64_t* part, int64_t n_parts, const float* layout, int dim,
                  double scaling, float* out);
int64_t orc_mst_boruvka(const float* emb, int64_t n, int dim, int ms, int32_t* u, int32_t* v,
                        float* w);
int orc_core_sq_kd(const float* emb, int64_t n, int dim, int ms, float* core);
// 0: dense Prim (the parity reference), 1: kd-tree Boruvka; returns the previous choice
int orc_set_mst_algo(int algo);

#ifdef __cplusplus
}
#endif
