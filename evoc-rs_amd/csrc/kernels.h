// Internal (C++) interface between the host orchestration and the HIP kernels.
#pragma once
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "common.h"

namespace evoc {

// ------------------------------------------------------------------ device memory
// Caching device allocator: freed blocks are kept by rounded size and reused, so a
// second call of the pipeline on the same shapes performs no hipMalloc.
class DevPool {
 public:
  void* get(size_t bytes) {
    size_t r = round(bytes);
    std::lock_guard<std::mutex> g(mu_);
    auto it = free_.find(r);
    if (it != free_.end()) {
      void* p = it->second;
      free_.erase(it);
      return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, r);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      release_locked();
      EVOC_HIP(hipMalloc(&p, r));
    }
    return p;
  }
  void put(void* p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu_);
    free_.insert({round(bytes), p});
  }
  void release() {
    std::lock_guard<std::mutex> g(mu_);
    release_locked();
  }
  // One pinned host staging buffer of at least `bytes` (grown on demand, contents not
  // kept), for device-to-host copies at full link speed; one user at a time.
  void* pinned(size_t bytes) {
    std::lock_guard<std::mutex> g(mu_);
    if (bytes > pinned_bytes_) {
      if (pinned_) (void)hipHostFree(pinned_);
      pinned_ = nullptr;
      pinned_bytes_ = 0;
      EVOC_HIP(hipHostMalloc(&pinned_, bytes, hipHostMallocDefault));
      pinned_bytes_ = bytes;
    }
    return pinned_;
  }
  ~DevPool() { release(); }

 private:
  static size_t round(size_t b) {
    if (b < 256) return 256;
    if (b < (1u << 20)) {
      size_t r = 256;
      while (r < b) r <<= 1;
      return r;
    }
    const size_t g = 2u << 20;
    return (b + g - 1) / g * g;
  }
  void release_locked() {
    for (auto& kv : free_) (void)hipFree(kv.second);
    free_.clear();
    if (pinned_) (void)hipHostFree(pinned_);
    pinned_ = nullptr;
    pinned_bytes_ = 0;
  }
  std::mutex mu_;
  std::multimap<size_t, void*> free_;
  void* pinned_ = nullptr;
  size_t pinned_bytes_ = 0;
};

template <typename T>
struct Buf {
  T* p = nullptr;
  size_t n = 0;
  DevPool* pool = nullptr;
  Buf() = default;
  Buf(DevPool& pl, size_t count) { alloc(pl, count); }
  void alloc(DevPool& pl, size_t count) {
    reset();
    pool = &pl;
    n = count;
    p = static_cast<T*>(pl.get(std::max<size_t>(count, 1) * sizeof(T)));
  }
  void reset() {
    if (p && pool) pool->put(p, std::max<size_t>(n, 1) * sizeof(T));
    p = nullptr;
    n = 0;
  }
  ~Buf() { reset(); }
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  Buf(Buf&& o) noexcept : p(o.p), n(o.n), pool(o.pool) { o.p = nullptr; o.n = 0; }
  Buf& operator=(Buf&& o) noexcept {
    if (this != &o) {
      reset();
      p = o.p; n = o.n; pool = o.pool;
      o.p = nullptr; o.n = 0;
    }
    return *this;
  }
};

// ------------------------------------------------------------------ kNN
enum KnnPrecision { KNN_AUTO = 0, KNN_SPLIT = 1, KNN_BF16 = 2, KNN_FP8 = 3 };
struct KnnStats {
  int split = 0;
  int fallback_rows = 0;
  int candidates = 0;
  float mfma_ms = 0.0f;
  double flops = 0.0;
  int symmetric = 0;  // the symmetric sweep ran
  int escalated_rows = 0;  // rows searched again with 64 candidates after a failed certificate
  int first_pass_failed = 0;  // rows whose first-pass certificate failed
};
// The first tile pass's raw candidate lists (two column halves of C/2 per row, keys in the
// pass's own units) and the per-row key scales, copied to host memory when requested
// (the stage entry point evoc_hip_knn_tile_pass, for tests of the tile kernels).
struct KnnCapture {
  std::vector<int32_t> idx;
  std::vector<float> key;
  std::vector<float> scale;
  int C = 0;
  float gamma = 0.0f;  // the certificate's relative dot error term of the pass
};
// the calling thread's last knn_device call
KnnStats& last_knn_stats();
// verbose > 0 prints the tile pass, the escalation and the exact search as they finish
void knn_device(const float* dX, int64_t n, int d, int64_t ld, int metric, int k, int precision,
                int32_t* d_idx, float* d_dist, KnnStats* stats, hipStream_t s, DevPool& pool,
                int verbose = 0, KnnCapture* capture = nullptr);

// ------------------------------------------------------------------ fuzzy set
struct Graph {  // CSR, rows sorted by column when symmetrised
  Buf<int64_t> indptr;
  Buf<int32_t> indices;
  Buf<float> weights;
  int64_t n = 0, nnz = 0;
};
void fuzzy_device(const int32_t* d_idx, const float* d_dist, int64_t n, int k, float effective_k,
                  bool symmetrise, float* d_sigma, float* d_rho, Graph& out, hipStream_t s,
                  DevPool& pool);

// ------------------------------------------------------------------ embedding
struct EmbedParams {
  int n_epochs = 50;
  float noise_level = 0.5f;
  float negative_sample_rate = 1.0f;
  float initial_alpha = 0.1f;
  float beta1 = 0.5f, beta2 = 0.9f, eps = 1e-7f;
  // nodes with at least this many edges take the wavefront-per-node epoch kernel (same
  // arithmetic, so any value gives the same layout); small coarse graphs use 1
  int64_t wide_deg = 64;
};
// d_emb holds the initial layout on entry (or is filled from Philox when
// random_init) and the final layout on exit.
void embed_device(const Graph& g, int dim, const EmbedParams& p, float* d_emb, bool random_init,
                  uint64_t seed, hipStream_t s, DevPool& pool, float* kernel_ms = nullptr);

// ------------------------------------------------------------------ label propagation
// rows of (R^T_norm)_l1 X for a partition given as ascending member lists:
// out[a] = coef[a] * sum over members m of a (in list order) of X[m]; host in/out
void partition_means_device(const float* dX, int64_t ld, int64_t d, const int64_t* h_mptr,
                            const int32_t* h_mem, const float* h_coef, int64_t np, int64_t n,
                            float* h_out, hipStream_t s, DevPool& pool);

// partition means with the member lists already in device memory (out: np x d)
void partition_means_dev(const float* dX, int64_t ld, int64_t d, const int64_t* d_mptr,
                         const int32_t* d_mem, const float* d_coef, int64_t np, float* d_out,
                         hipStream_t s);
// the input level's partition (label propagation, outliers, remap) in device memory
struct LpTop {
  Buf<int32_t> part;  // partition of every node
  Buf<int32_t> mem;   // nodes by partition, ascending within one
  Buf<int64_t> mptr;  // np + 1 offsets into mem
  Buf<float> sa;      // 1/sqrt(|a|): column norms of the partition map
  Buf<float> coef;    // L1 row normalisation of R^T_norm (partition means)
  int64_t np = 0;
};
void lp_top_partition(const int64_t* d_ptr, const int32_t* d_col, const float* d_val, int64_t n,
                      int64_t approx, int n_iter, uint64_t seed, LpTop& P, hipStream_t s,
                      DevPool& pool);
// R^T_norm G R of a level, returned as a host CSR (np x np) and, with keep, also left in
// device memory (then host_copy = false skips the host CSR of a non-empty graph)
void lp_top_coarsen(const int64_t* d_ptr, const int32_t* d_col, const float* d_val, int64_t n,
                    int64_t nnz, const LpTop& P, std::vector<int64_t>& cptr,
                    std::vector<int32_t>& ccol, std::vector<float>& cval, hipStream_t s,
                    DevPool& pool, Graph* keep = nullptr, bool host_copy = true);
// the embedding slots of a coarse graph in device memory (see coarse_embedding)
void coarse_slots_device(const Graph& g, Graph& out, hipStream_t s, DevPool& pool);
// expansion of the input level, centred and scaled, into d_out (n x dim)
void expansion_device(const int64_t* d_indptr, const int32_t* d_col, const float* d_val,
                      int64_t n, int64_t nnz, int dim, const int32_t* d_part, const float* d_sa,
                      int64_t np, const float* h_layout, float* d_out, hipStream_t s,
                      DevPool& pool);

// ------------------------------------------------------------------ core distances + MST
void core_sq_device(const float* d_emb, int64_t n, int dim, int min_samples, float* d_core,
                    hipStream_t s, DevPool& pool);
// writes n-1 edges (u < v) sorted by (w, u, v); weights sqrt'd
void mst_device(const float* d_emb, int64_t n, int dim, int min_samples, int32_t* h_u,
                int32_t* h_v, float* h_w, hipStream_t s, DevPool& pool, int* rounds = nullptr);

// ------------------------------------------------------------------ condensed tree
// The single-linkage tree (rows as src/clustering/linkage.rs:301-310: row i is node n + i,
// the root is node 2n - 2) in device memory, with every node's parent, its position in the
// right-first pre-order walk of condense_tree and the number of points before it there.
struct DevTree {
  int64_t n = 0, nr = 0;
  Buf<int32_t> rows;  // left, right, size: nr each
  Buf<float> dist;
  Buf<int32_t> par, pre, lf;  // 2n - 1 nodes each
  const int32_t* left() const { return rows.p; }
  const int32_t* right() const { return rows.p + nr; }
  const int32_t* size() const { return rows.p + 2 * nr; }
};
void tree_upload(const int64_t* left, const int64_t* right, const float* dist,
                 const int64_t* size, int64_t n, DevTree& t, hipStream_t s, DevPool& pool);
// rows of the condensed tree in the reference's order (src/clustering/condensed_tree.rs:19-37)
struct CondensedRows {
  std::vector<int64_t> parent, child, size;
  std::vector<float> lambda;
};
// One condensation. For min_cluster_size >= 2 every point falls out of exactly one
// cluster: pcl/plam hold, per point, that cluster and lambda (device). Split s (in the
// walk's order) opens clusters n + 1 + 2s (left child, split_ls points) and n + 2 + 2s
// (right child) under cluster split_parent[s] at split_lam[s] (host). S splits, E rows.
struct DevCondensed {
  int64_t n = 0, mcs = 0, S = 0, E = 0;
  Buf<int32_t> pcl;
  Buf<float> plam;
  std::vector<int32_t> split_parent, split_ls, split_rs;
  std::vector<float> split_lam;
};
// full: also every row, in the reference's order, copied to the host
void tree_condense(const DevTree& t, int64_t mcs, DevCondensed& out, CondensedRows* full,
                   hipStream_t s, DevPool& pool);
// Labels and membership strengths of nl layers of one condensation (min_cluster_size >= 2).
// Per layer l and cluster c (index c - n, C = 2S + 1 of them): clab[l*C + c] the label of
// a point that falls out of c (-1: none), dsel[l*C + c] the rank of c among the layer's
// nsel[l] selected clusters (-1: not selected). Per layer: ncl[l] clusters (largest label
// + 1) and assigned[l] labelled points. The layers are copied to the host in descending
// order of ncl (a stable sort, order[q] = the layer at position q): labels (int64) and
// strengths of order[q] to labels + q*n and strengths + q*n.
void tree_layer_labels(const DevCondensed& c, int nl, const std::vector<int32_t>& clab,
                       const std::vector<int32_t>& dsel, const std::vector<int64_t>& nsel,
                       int64_t* labels, float* strengths, std::vector<int64_t>& order,
                       std::vector<int64_t>& ncl, std::vector<int64_t>& assigned, hipStream_t s,
                       DevPool& pool);
// memcpy by several host threads (large copies into freshly allocated host memory)
void parallel_copy(void* dst, const void* src, size_t bytes);
// positions of layers sorted by descending cluster count, ties in layer order
inline std::vector<int64_t> order_by_clusters(const std::vector<int64_t>& ncl) {
  std::vector<int64_t> order(ncl.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int64_t)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return ncl[b] < ncl[a]; });
  return order;
}

// ------------------------------------------------------------------ sort
// stable LSD radix sort of (key, value) pairs on the low key_bits bits of key
void radix_sort_pairs(uint64_t* d_key, uint32_t* d_val, int64_t n, int key_bits, hipStream_t s,
                      DevPool& pool);

// ------------------------------------------------------------------ input checks
// throws std::invalid_argument("<what> contains a non-finite value") if any of the n x d
// row-major values (leading dimension ld) is NaN or +-inf
void check_finite_device(const float* dX, int64_t n, int d, int64_t ld, const char* what,
                         hipStream_t s, DevPool& pool);
inline void check_finite_host(const float* p, int64_t count, const char* what) {
  bool bad = false;
#pragma omp parallel for reduction(|| : bad) schedule(static) if (count > (1 << 20))
  for (int64_t i = 0; i < count; ++i) bad = bad || !std::isfinite(p[i]);
  EVOC_CHECK(!bad, std::string(what) + " contains a non-finite value");
}

// ------------------------------------------------------------------ scans
void exclusive_scan_i64(const int64_t* d_in, int64_t* d_out, int64_t n, hipStream_t s,
                        DevPool& pool);

}  // namespace evoc
