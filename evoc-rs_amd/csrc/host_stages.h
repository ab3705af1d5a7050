// Host-side stages of the pipeline (sequential tree logic and the label-propagation
// initial layout) and their device helpers.
#pragma once
#include <vector>

#include "kernels.h"

namespace evoc {

// data_dev: the n x d input rows in device memory (row stride ld floats), or null for
// the random base case
void label_prop_init_host(const int64_t* indptr, const int32_t* indices, const float* weights,
                          int64_t n, int dim, const float* data_dev, int64_t d, int64_t ld,
                          uint64_t seed, float* out, hipStream_t s, DevPool& pool, int verbose);
// same, with the graph in device memory (rows column-sorted) and the layout written to
// d_out (n x dim): the input level runs on the GPU, the coarse levels on the host
void label_prop_init_device(const int64_t* d_indptr, const int32_t* d_indices,
                            const float* d_weights, int64_t nnz, int64_t n, int dim,
                            const float* data_dev, int64_t d, int64_t ld, uint64_t seed,
                            float* d_out, hipStream_t s, DevPool& pool, int verbose);

// scipy-style single-linkage rows (src/clustering/linkage.rs:301-310)
struct Linkage {
  std::vector<int64_t> left, right, size;
  std::vector<float> dist;
  int64_t n = 0;
};
// edges must already be in (w, u, v) order
Linkage linkage_from_sorted(const int32_t* u, const int32_t* v, const float* w, int64_t ne,
                            int64_t n);
// the same into L, whose vectors keep their capacity from call to call
void linkage_from_sorted(const int32_t* u, const int32_t* v, const float* w, int64_t ne,
                         int64_t n, Linkage& L);

// condensed tree rows (src/clustering/condensed_tree.rs:19-37)
struct CTree {
  std::vector<int64_t> parent, child, size;
  std::vector<float> lambda;
  int64_t n = 0;
};
std::vector<int64_t> leaves(const CTree& t);
void labels_and_strengths(const CTree& t, const std::vector<int64_t>& clusters, int64_t* labels,
                          float* strengths);

// cluster-level views of a device condensation (min_cluster_size >= 2)
CTree split_rows(const DevCondensed& c);
std::vector<int64_t> split_leaves(const DevCondensed& c);
void layer_tables(const DevCondensed& c, const std::vector<std::vector<int64_t>>& sel,
                  std::vector<int32_t>& clab, std::vector<int32_t>& dsel,
                  std::vector<int64_t>& nsel);
// the approx_n_clusters binary search's final bracket [lo, hi] of min_cluster_size
void search_bounds(const Linkage& L, int64_t target_k, int64_t& lo, int64_t& hi);

// The layers of a clustering in host memory: labels and strengths nl x n (malloc'd, so a
// result can take them over), in descending order of cluster count, and each layer's
// persistence.
struct Layers {
  int64_t n = 0;
  int64_t* labels = nullptr;
  float* strengths = nullptr;
  std::vector<double> persistence;
  Layers() = default;
  Layers(int64_t n_, int nl);
  ~Layers();
  Layers(Layers&& o) noexcept;
  Layers& operator=(Layers&& o) noexcept;
  Layers(const Layers&) = delete;
  Layers& operator=(const Layers&) = delete;
  int nl() const { return (int)persistence.size(); }
  const int64_t* layer(int l) const { return labels + (size_t)l * n; }
};
// the selected clusters of every layer and its persistence, from rows holding at least
// every cluster row (size > 1) in the condensed tree's order, given the base layer's leaves
struct LayerSelection {
  std::vector<std::vector<int64_t>> sel;
  std::vector<double> pers;
};
LayerSelection select_layers(const CTree& t, std::vector<int64_t> base, int64_t base_mcs,
                             double min_sim, int64_t max_layers);
// layers given in selection order (nl x n) put in descending order of cluster count
Layers order_layers(int64_t n, const int64_t* lab, const float* str,
                    const std::vector<int64_t>& ncl, const std::vector<double>& pers);
Layers layers_from_rows(const CTree& t, const LayerSelection& s);

// the tree stages on the device (clusters.cpp): condensation, layers, the n-cluster search
CTree condense_device(const Linkage& L, int64_t mcs, hipStream_t s, DevPool& pool);
Layers cluster_layers_device(const Linkage& L, int64_t base_mcs, double min_sim,
                             int64_t max_layers, hipStream_t s, DevPool& pool, int verbose = 0);
void search_n_clusters_device(const Linkage& L, int64_t target_k, int64_t* labels,
                              float* strengths, hipStream_t s, DevPool& pool);
int best_layer_index(const std::vector<double>& persistence);
// adjusted Rand index of two labellings of n points (-1 counts as a label)
double adjusted_rand_index(const int64_t* a, const int64_t* b, int64_t n);

}  // namespace evoc
