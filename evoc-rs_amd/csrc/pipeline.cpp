// Pipeline orchestration (evoc(), src/lib.rs:199-352) and the stage entry points.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <thread>
#include <unordered_set>

#include "api_internal.h"
#include "host_stages.h"
#include "mlp.h"

namespace evoc {
namespace {

using clk = std::chrono::steady_clock;
double ms_since(clk::time_point t) {
  return std::chrono::duration<double, std::milli>(clk::now() - t).count();
}

template <typename T>
void to_host(std::vector<T>& h, const T* d, size_t n, hipStream_t s) {
  h.resize(n);
  if (n) EVOC_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s));
}
template <typename T>
void to_dev(Buf<T>& b, const T* h, size_t n, DevPool& pool, hipStream_t s) {
  b.alloc(pool, n);
  if (n) EVOC_HIP(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
}

// Embedding slots for an arbitrary adjacency (evoc_embedding, src/graph/embedding.rs:79-135):
// edges (i, j, w) with i < j are taken from row i in row order; each endpoint gets one slot,
// slots of a node ordered by edge index. For a symmetric graph with column-sorted rows
// this is the graph itself.
void build_slots(const std::vector<int64_t>& ptr, const std::vector<int32_t>& col,
                 const std::vector<float>& w, std::vector<int64_t>& sp, std::vector<int32_t>& so,
                 std::vector<float>& sw) {
  const int64_t n = (int64_t)ptr.size() - 1;
  std::vector<int64_t> deg(n, 0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p)
      if (i < col[p]) {
        deg[i]++;
        deg[col[p]]++;
      }
  sp.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) sp[i + 1] = sp[i] + deg[i];
  so.resize(sp[n]);
  sw.resize(sp[n]);
  std::vector<int64_t> cur(sp.begin(), sp.end() - 1);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) {
      const int32_t j = col[p];
      if (i < j) {
        so[cur[i]] = j;
        sw[cur[i]++] = w[p];
        so[cur[j]] = (int32_t)i;
        sw[cur[j]++] = w[p];
      }
    }
}

bool symmetric_sorted(const std::vector<int64_t>& ptr, const std::vector<int32_t>& col) {
  const int64_t n = (int64_t)ptr.size() - 1;
  for (int64_t i = 0; i < n; ++i)
    for (int64_t p = ptr[i]; p < ptr[i + 1]; ++p) {
      if (col[p] == i || col[p] < 0 || col[p] >= n) return false;
      if (p > ptr[i] && col[p] <= col[p - 1]) return false;
      const int32_t* b = col.data() + ptr[col[p]];
      const int32_t* e = col.data() + ptr[col[p] + 1];
      const int32_t* f = std::lower_bound(b, e, (int32_t)i);
      if (f == e || *f != i) return false;
    }
  return true;
}

int auto_dim(const evoc_params_t& p) {
  return p.embedding_dim > 0 ? p.embedding_dim : std::min(std::max(p.n_neighbours / 4, 4), 16);
}

void validate_params(const evoc_params_t& p) {
  EVOC_CHECK(p.n_neighbours >= 1 && p.n_neighbours <= 64, "n_neighbours must be in [1, 64]");
  EVOC_CHECK(p.n_epochs >= 0, "n_epochs must be >= 0");
  EVOC_CHECK(p.embedding_dim >= 0 && p.embedding_dim <= 16, "embedding_dim must be in [0, 16]");
  EVOC_CHECK(p.min_samples >= 0 && p.min_samples <= 64, "min_samples must be in [0, 64]");
  EVOC_CHECK(p.base_min_cluster_size >= 0, "base_min_cluster_size must be >= 0");
  EVOC_CHECK(p.approx_n_clusters >= 0, "approx_n_clusters must be >= 0");
  EVOC_CHECK(p.max_layers >= 0, "max_layers must be >= 0");
}

void free_result(evoc_result_t* r);

// Device-to-host copy of finished results on the context's second stream, driven by a
// helper thread (a copy into pageable memory holds its calling thread until it is done),
// so that the later stages run meanwhile. The destructor waits for it.
struct BackCopy {
  std::thread t;
  hipError_t err = hipSuccess;
  void start(Ctx& c, void* d1, const void* s1, size_t b1, void* d2, const void* s2, size_t b2) {
    t = std::thread([this, &c, d1, s1, b1, d2, s2, b2]() {
      hipError_t e = hipSetDevice(c.device);
      if (e == hipSuccess && b1) e = hipMemcpyAsync(d1, s1, b1, hipMemcpyDeviceToHost, c.aux);
      if (e == hipSuccess && b2) e = hipMemcpyAsync(d2, s2, b2, hipMemcpyDeviceToHost, c.aux);
      if (e == hipSuccess) e = hipStreamSynchronize(c.aux);
      err = e;
    });
  }
  void finish() {
    if (t.joinable()) t.join();
    EVOC_HIP(err);
  }
  ~BackCopy() {
    if (t.joinable()) t.join();
  }
};

// host buffers of a result, freed unless handed over to it
struct HostOut {
  int32_t* idx = nullptr;
  float* dist = nullptr;
  float* emb = nullptr;
  ~HostOut() {
    std::free(idx);
    std::free(dist);
    std::free(emb);
  }
};
template <typename T>
T* host_alloc(size_t count) {
  void* p = std::malloc(count ? count * sizeof(T) : 1);
  if (!p) throw std::bad_alloc();
  return static_cast<T*>(p);
}

evoc_result_t* alloc_result(int64_t n, int k, int dim, int nl, HostOut& take, Layers& lay) {
  auto* r = static_cast<evoc_result_t*>(std::calloc(1, sizeof(evoc_result_t)));
  if (!r) throw std::bad_alloc();
  r->n = n;
  r->k = k;
  r->dim = dim;
  r->n_layers = nl;
  auto al = [](size_t bytes) {
    void* p = std::malloc(bytes ? bytes : 1);
    if (!p) throw std::bad_alloc();
    return p;
  };
  // the kNN graph and the embedding: buffers already filled, handed over
  std::swap(r->nn_indices, take.idx);
  std::swap(r->nn_distances, take.dist);
  std::swap(r->embedding, take.emb);
  // the layers' labels and strengths likewise (malloc'd nl x n arrays)
  std::swap(r->cluster_layers, lay.labels);
  std::swap(r->membership_strengths, lay.strengths);
  try {
    r->persistence_scores = static_cast<double*>(al((size_t)nl * sizeof(double)));
  } catch (...) {
    free_result(r);
    throw;
  }
  for (int l = 0; l < nl; ++l) r->persistence_scores[l] = lay.persistence[l];
  return r;
}

void free_result(evoc_result_t* r) {
  if (!r) return;
  std::free(r->cluster_layers);
  std::free(r->membership_strengths);
  std::free(r->persistence_scores);
  std::free(r->nn_indices);
  std::free(r->nn_distances);
  std::free(r->embedding);
  std::free(r);
}

// A host matrix with any element strides (faer's column-major MatRef included) into a
// row-major device copy, through the pool's pinned staging buffer: host threads pack one
// 32 MB chunk of rows into one half while the other half's copy crosses the link, so the
// packing overlaps the transfer and the transfer runs at pinned-memory speed.
void upload_rows(const float* data, int64_t n, int64_t d, int64_t rs, int64_t cs, float* dst,
                 Ctx& c) {
  constexpr int64_t CHUNK = int64_t(32) << 20;  // bytes per half
  const int64_t per = std::max<int64_t>(1, CHUNK / (d * (int64_t)sizeof(float)));
  float* stage = static_cast<float*>(c.pool.pinned((size_t)(2 * per * d) * sizeof(float)));
  hipEvent_t ev[2];
  EVOC_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  EVOC_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  bool pending[2] = {false, false};
  hipError_t err = hipSuccess;
  int h = 0;
  for (int64_t r0 = 0; r0 < n && err == hipSuccess; r0 += per, h ^= 1) {
    const int64_t rows = std::min(per, n - r0);
    if (pending[h]) err = hipEventSynchronize(ev[h]);
    if (err != hipSuccess) break;
    float* buf = stage + h * per * d;
    if (rs == d && cs == 1) {
      parallel_copy(buf, data + r0 * d, (size_t)(rows * d) * sizeof(float));
    } else {
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < rows; ++i)
        for (int64_t j = 0; j < d; ++j) buf[i * d + j] = data[(r0 + i) * rs + j * cs];
    }
    err = hipMemcpyAsync(dst + r0 * d, buf, (size_t)(rows * d) * sizeof(float),
                         hipMemcpyHostToDevice, c.stream);
    if (err == hipSuccess) err = hipEventRecord(ev[h], c.stream);
    pending[h] = true;
  }
  const hipError_t e2 = hipStreamSynchronize(c.stream);
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  EVOC_HIP(err);
  EVOC_HIP(e2);
}

// The clustering stage on a device embedding: MST -> linkage -> layers or search.
Layers cluster_device(const float* d_emb, int64_t n, int dim, const evoc_params_t& p,
                      hipStream_t s, DevPool& pool, double* mst_ms, double* layer_ms,
                      int verbose = 0, double* mst_total = nullptr) {
  Layers out;
  auto t0 = clk::now();
  auto tl = t0;
  auto lap = [&](const char* what) {
    if (verbose >= 2) {
      std::printf("    %s: %.1f ms\n", what, ms_since(tl));
      tl = clk::now();
    }
  };
  // host arrays of the edges and the linkage, kept by the calling thread between calls (a
  // fresh allocation of these sizes is paid in page faults on every call)
  static thread_local std::vector<int32_t> u, v;
  static thread_local std::vector<float> w;
  static thread_local Linkage L;
  u.resize(std::max<int64_t>(n - 1, 1));
  v.resize(std::max<int64_t>(n - 1, 1));
  w.resize(std::max<int64_t>(n - 1, 1));
  mst_device(d_emb, n, dim, p.min_samples, u.data(), v.data(), w.data(), s, pool);
  lap("mst");
  int64_t ne = std::max<int64_t>(n - 1, 0);
  while (ne > 0 && u[ne - 1] < 0) --ne;
  // A spanning tree is guaranteed for finite embeddings whose squared distances do not
  // overflow (every entry point rejects non-finite input); the condensed tree is rooted
  // at the last linkage node, so a forest must not reach it.
  EVOC_CHECK(ne == std::max<int64_t>(n - 1, 0),
             "the minimum spanning tree is disconnected (embedding distances overflow f32)");
  if (mst_total) {
    double t = 0.0;
    for (int64_t e = 0; e < ne; ++e) t += (double)w[e];
    *mst_total = t;
  }
  linkage_from_sorted(u.data(), v.data(), w.data(), ne, n, L);
  lap("linkage");
  if (mst_ms) *mst_ms = ms_since(t0);
  t0 = clk::now();
  if (p.approx_n_clusters > 0) {
    out = Layers(n, 1);
    search_n_clusters_device(L, p.approx_n_clusters, out.labels, out.strengths, s, pool);
  } else {
    out = cluster_layers_device(L, p.base_min_cluster_size, p.min_similarity_threshold,
                                p.max_layers, s, pool, verbose);
  }
  lap("layers");
  if (layer_ms) *layer_ms = ms_since(t0);
  return out;
}

evoc_result_t* run_impl(const float* dX, const float* hX, int64_t n, int64_t d, int64_t ld,
                        const int32_t* pre_idx, const float* pre_dist, int pre_k,
                        const evoc_params_t& p, const evoc_hip_nn_params_t& nn, uint64_t seed,
                        int verbose, Ctx& c) {
  validate_params(p);
  hipStream_t s = c.stream;
  DevPool& pool = c.pool;
  const auto t_all = clk::now();
  check_finite_device(dX, n, (int)d, ld, "input", s, pool);
  double stage[8] = {0};
  const int k = pre_idx ? pre_k : p.n_neighbours;
  // 1. kNN
  auto t0 = clk::now();
  Buf<int32_t> didx(pool, n * k);
  Buf<float> ddist(pool, n * k);
  int64_t fallback = 0;
  double knn_kernel_ms = 0.0;
  float embed_kernel_ms = 0.0f;
  if (pre_idx) {
    if (verbose) std::printf("Using precomputed kNN graph...\n");
    for (int64_t i = 0; i < n; ++i) {
      std::unordered_set<int32_t> seen;
      for (int t = 0; t < k; ++t) {
        const int32_t j = pre_idx[i * k + t];
        const float dj = pre_dist[i * k + t];
        EVOC_CHECK(j >= -1 && j < n, "precomputed kNN index out of range");
        EVOC_CHECK(j < 0 || seen.insert(j).second, "precomputed kNN row repeats an index");
        EVOC_CHECK(j < 0 || (std::isfinite(dj) && dj >= 0.0f),
                   "precomputed kNN distance is negative or non-finite");
      }
    }
    EVOC_HIP(hipMemcpyAsync(didx.p, pre_idx, n * k * sizeof(int32_t), hipMemcpyHostToDevice, s));
    EVOC_HIP(hipMemcpyAsync(ddist.p, pre_dist, n * k * sizeof(float), hipMemcpyHostToDevice, s));
  } else {
    if (verbose) std::printf("Running exhaustive nearest neighbour search on the GPU...\n");
    KnnStats st;
    knn_device(dX, n, (int)d, ld, nn.metric, k, nn.knn_precision, didx.p, ddist.p, &st, s, pool,
               verbose);
    fallback = st.fallback_rows;
    knn_kernel_ms = st.mfma_ms;
  }
  EVOC_HIP(hipStreamSynchronize(s));
  stage[0] = ms_since(t0);
  if (verbose) std::printf("kNN search done in %.2f ms.\n", stage[0]);
  HostOut ho;
  ho.idx = host_alloc<int32_t>((size_t)n * k);
  ho.dist = host_alloc<float>((size_t)n * k);
  BackCopy knn_back;  // joined before didx / ddist are released
  knn_back.start(c, ho.idx, didx.p, (size_t)n * k * sizeof(int32_t), ho.dist, ddist.p,
                 (size_t)n * k * sizeof(float));
  // 2. fuzzy simplicial set
  t0 = clk::now();
  if (verbose) std::printf("Constructing fuzzy simplicial set...\n");
  const float eff_k = p.neighbour_scale * (float)p.n_neighbours;
  Graph g;
  fuzzy_device(didx.p, ddist.p, n, k, eff_k, p.symmetrise != 0, nullptr, nullptr, g, s, pool);
  EVOC_HIP(hipStreamSynchronize(s));
  stage[1] = ms_since(t0);
  if (verbose) std::printf("... fuzzy simplicial set done in %.2f ms.\n", stage[1]);
  const int dim = auto_dim(p);
  // graph on the host: needed by the non-symmetrised initial layout and embedding
  std::vector<int64_t> hptr;
  std::vector<int32_t> hcol;
  std::vector<float> hw;
  // 3. initial layout
  t0 = clk::now();
  Buf<float> demb(pool, n * dim);
  int epochs = p.n_epochs;
  MlpWarmStart ws;
  bool warm_used = false;
  if (nn.warm_start) {
    if (verbose) std::printf("Computing learned warm-start initialisation...\n");
    ws.subsample = nn.mlp_subsample > 0 ? nn.mlp_subsample : 50000;
    const int warm_epochs =
        nn.warm_epochs > 0 ? nn.warm_epochs : std::max(1, (int)(0.3 * p.n_epochs));
    // The learned layout is accepted only if, on the training subsample, the warm-started
    // short run reproduces the plain run: best-layer labels with ARI >= 0.98 and the MST
    // total weight within 1%. Otherwise the plain initial layout and epoch count are used.
    WarmJudge judge = [&](const Graph& sg, const float* d_plain, const float* d_pred, int64_t S,
                          MlpWarmStart& w) {
      Buf<float> dw(pool, (size_t)S * dim);
      EVOC_HIP(hipMemcpyAsync(dw.p, d_pred, (size_t)S * dim * sizeof(float),
                              hipMemcpyDeviceToDevice, s));
      EmbedParams wp;
      wp.n_epochs = warm_epochs;
      wp.noise_level = p.noise_level;
      wp.initial_alpha = 0.1f;
      embed_device(sg, dim, wp, dw.p, false, seed, s, pool);
      EVOC_HIP(hipStreamSynchronize(s));
      double mst_plain = 0.0, mst_warm = 0.0;
      Layers lp = cluster_device(d_plain, S, dim, p, s, pool, nullptr, nullptr, 0, &mst_plain);
      Layers lw = cluster_device(dw.p, S, dim, p, s, pool, nullptr, nullptr, 0, &mst_warm);
      w.judge_ari = adjusted_rand_index(lp.layer(best_layer_index(lp.persistence)),
                                        lw.layer(best_layer_index(lw.persistence)), S);
      w.judge_mst_ratio = mst_plain > 0.0 ? mst_warm / mst_plain : 1.0;
      return w.judge_ari >= 0.98 && std::fabs(w.judge_mst_ratio - 1.0) <= 0.01;
    };
    warm_used = warm_start_init(dX, n, (int)d, ld, p, nn, dim, seed, ws, judge, demb.p, s, pool,
                                verbose);
    if (warm_used) epochs = warm_epochs;
  }
  if (!warm_used) {
    if (verbose) std::printf("Computing label propagation initialisation...\n");
    if (p.symmetrise) {  // the device CSR is the column-sorted graph
      label_prop_init_device(g.indptr.p, g.indices.p, g.weights.p, g.nnz, n, dim, dX, d, ld, seed,
                             demb.p, s, pool, verbose);
    } else {
      to_host(hptr, g.indptr.p, n + 1, s);
      to_host(hcol, g.indices.p, g.nnz, s);
      to_host(hw, g.weights.p, g.nnz, s);
      EVOC_HIP(hipStreamSynchronize(s));
      std::vector<float> init(n * dim);
      // Csr::from_coo (src/utils/sparse.rs:195-235) orders every row by column
      std::vector<int32_t> c2(hcol);
      std::vector<float> w2(hw);
      for (int64_t i = 0; i < n; ++i) {
        std::vector<std::pair<int32_t, float>> row;
        for (int64_t q = hptr[i]; q < hptr[i + 1]; ++q) row.push_back({hcol[q], hw[q]});
        std::sort(row.begin(), row.end(), [](const std::pair<int32_t, float>& a,
                                             const std::pair<int32_t, float>& b) {
          return a.first < b.first;
        });
        for (size_t q = 0; q < row.size(); ++q) {
          c2[hptr[i] + q] = row[q].first;
          w2[hptr[i] + q] = row[q].second;
        }
      }
      label_prop_init_host(hptr.data(), c2.data(), w2.data(), n, dim, dX, d, ld, seed,
                           init.data(), s, pool, verbose);
      EVOC_HIP(hipMemcpyAsync(demb.p, init.data(), n * dim * sizeof(float), hipMemcpyHostToDevice,
                              s));
    }
  }
  EVOC_HIP(hipStreamSynchronize(s));
  stage[2] = ms_since(t0);
  if (verbose) std::printf("Label prop init done in %.2f ms.\n", stage[2]);
  // 4. node embedding
  t0 = clk::now();
  if (verbose) std::printf("Computing %d-d node embedding (%d epochs)...\n", dim, epochs);
  EmbedParams ep;
  ep.n_epochs = epochs;
  ep.noise_level = p.noise_level;
  ep.initial_alpha = 0.1f;
  if (p.symmetrise) {
    embed_device(g, dim, ep, demb.p, false, seed, s, pool, &embed_kernel_ms);
  } else {
    if (hptr.empty()) {
      to_host(hptr, g.indptr.p, n + 1, s);
      to_host(hcol, g.indices.p, g.nnz, s);
      to_host(hw, g.weights.p, g.nnz, s);
      EVOC_HIP(hipStreamSynchronize(s));
    }
    std::vector<int64_t> sp;
    std::vector<int32_t> so;
    std::vector<float> sw;
    build_slots(hptr, hcol, hw, sp, so, sw);
    Graph eg;
    eg.n = n;
    eg.nnz = sp[n];
    to_dev(eg.indptr, sp.data(), n + 1, pool, s);
    to_dev(eg.indices, so.data(), eg.nnz, pool, s);
    to_dev(eg.weights, sw.data(), eg.nnz, pool, s);
    embed_device(eg, dim, ep, demb.p, false, seed, s, pool, &embed_kernel_ms);
  }
  EVOC_HIP(hipStreamSynchronize(s));
  stage[3] = ms_since(t0);
  ho.emb = host_alloc<float>((size_t)n * dim);
  BackCopy emb_back;  // joined before demb is released
  emb_back.start(c, ho.emb, demb.p, (size_t)n * dim * sizeof(float), nullptr, nullptr, 0);
  if (verbose) std::printf(" ... embedding done in %.2f ms.\n", stage[3]);
  // 5./6. clustering
  if (verbose) std::printf("Running density-based clustering...\n");
  Layers lay = cluster_device(demb.p, n, dim, p, s, pool, &stage[4], &stage[5], verbose);
  const int nl = lay.nl();
  knn_back.finish();
  emb_back.finish();
  evoc_result_t* r = alloc_result(n, k, dim, nl, ho, lay);
  r->best_layer = best_layer_index(lay.persistence);
  r->knn_fallback_rows = fallback;
  r->warm_start_used = warm_used ? 1 : 0;
  r->warm_start_ari = ws.judge_ari;
  r->warm_start_mst_ratio = ws.judge_mst_ratio;
  r->knn_kernel_ms = knn_kernel_ms;
  r->embed_kernel_ms = embed_kernel_ms;
  stage[6] = ms_since(t_all);
  for (int i = 0; i < 8; ++i) r->stage_ms[i] = stage[i];
  if (verbose) {
    std::printf("Clustering done in %.2f ms: %d layer(s).\n", stage[4] + stage[5], nl);
    std::printf("EVoC (HIP) total: %.2f ms.\n", stage[6]);
  }
  return r;
}

}  // namespace
}  // namespace evoc

using namespace evoc;

extern "C" {

int evoc_hip_run(const float* data, int64_t n, int64_t d, int64_t row_stride, int64_t col_stride,
                 const int32_t* pre_idx, const float* pre_dist, int32_t pre_k,
                 const evoc_params_t* params, const evoc_hip_nn_params_t* nn, uint64_t seed,
                 int32_t verbose, evoc_result_t** out) {
  return guarded([&] {
    EVOC_CHECK(out, "out is NULL");
    *out = nullptr;
    EVOC_CHECK(data && params, "data or params is NULL");
    EVOC_CHECK(n >= 2 && d >= 1, "need at least 2 points and 1 feature");
    EVOC_CHECK(!pre_idx == !pre_dist, "pre_idx and pre_dist must both be given");
    EVOC_CHECK(!pre_idx || (pre_k >= 1 && pre_k <= 64), "pre_k must be in [1, 64]");
    evoc_hip_nn_params_t q;
    if (nn) q = *nn;
    else evoc_hip_default_nn_params(&q);
    Ctx& c = context(q.device_id);
    const auto t0 = std::chrono::steady_clock::now();
    Buf<float> dx(c.pool, n * d);
    upload_rows(data, n, d, row_stride, col_stride, dx.p, c);
    const double up = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    evoc_result_t* r = run_impl(dx.p, nullptr, n, d, d, pre_idx, pre_dist, pre_k, *params, q, seed,
                                verbose, c);
    r->stage_ms[7] = up;
    *out = r;
  });
}

int evoc_hip_run_device(const float* d_data, int64_t n, int64_t d, int64_t ld,
                        const evoc_params_t* params, const evoc_hip_nn_params_t* nn,
                        uint64_t seed, int32_t verbose, evoc_result_t** out) {
  return guarded([&] {
    EVOC_CHECK(out, "out is NULL");
    *out = nullptr;
    EVOC_CHECK(d_data && params, "data or params is NULL");
    EVOC_CHECK(n >= 2 && d >= 1 && ld >= d, "need n >= 2, d >= 1, ld >= d");
    evoc_hip_nn_params_t q;
    if (nn) q = *nn;
    else evoc_hip_default_nn_params(&q);
    Ctx& c = context(q.device_id);
    *out = run_impl(d_data, nullptr, n, d, ld, nullptr, nullptr, 0, *params, q, seed, verbose, c);
  });
}

void evoc_hip_result_free(evoc_result_t* r) { free_result(r); }

int evoc_hip_fuzzy(const int32_t* knn_idx, const float* knn_dist, int64_t n, int32_t k,
                   float effective_k, int32_t symmetrise, float* sigma, float* rho,
                   int64_t* indptr, int32_t* indices, float* weights, int64_t* nnz) {
  return guarded([&] {
    EVOC_CHECK(knn_idx && knn_dist && indptr && indices && weights && nnz, "fuzzy: NULL buffer");
    EVOC_CHECK(n >= 1 && k >= 1 && k <= 64, "fuzzy: need n >= 1 and k in [1, 64]");
    check_finite_host(knn_dist, n * k, "knn_dist");
    Ctx& c = context(0);
    Buf<int32_t> di;
    Buf<float> dd, ds(c.pool, n), dr(c.pool, n);
    to_dev(di, knn_idx, n * k, c.pool, c.stream);
    to_dev(dd, knn_dist, n * k, c.pool, c.stream);
    Graph g;
    fuzzy_device(di.p, dd.p, n, k, effective_k, symmetrise != 0, ds.p, dr.p, g, c.stream, c.pool);
    if (sigma) EVOC_HIP(hipMemcpyAsync(sigma, ds.p, n * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    if (rho) EVOC_HIP(hipMemcpyAsync(rho, dr.p, n * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    EVOC_HIP(hipMemcpyAsync(indptr, g.indptr.p, (n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
    if (g.nnz) {
      EVOC_HIP(hipMemcpyAsync(indices, g.indices.p, g.nnz * sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
      EVOC_HIP(hipMemcpyAsync(weights, g.weights.p, g.nnz * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    }
    EVOC_HIP(hipStreamSynchronize(c.stream));
    *nnz = g.nnz;
  });
}

int evoc_hip_embed(const int64_t* indptr, const int32_t* indices, const float* weights, int64_t n,
                   int32_t dim, int32_t n_epochs, float noise_level, float negative_sample_rate,
                   float initial_alpha, float beta1, float beta2, float eps, const float* init,
                   uint64_t seed, float* out) {
  return guarded([&] {
    EVOC_CHECK(indptr && out, "embed: NULL buffer");
    EVOC_CHECK(n >= 0 && dim >= 1 && dim <= 16, "embed: dim must be in [1, 16]");
    if (n == 0) return;
    check_finite_host(weights, indptr[n], "weights");
    if (init) check_finite_host(init, n * dim, "init");
    Ctx& c = context(0);
    std::vector<int64_t> ptr(indptr, indptr + n + 1);
    std::vector<int32_t> col(indices, indices + ptr[n]);
    std::vector<float> w(weights, weights + ptr[n]);
    std::vector<int64_t> sp;
    std::vector<int32_t> so;
    std::vector<float> sw;
    build_slots(ptr, col, w, sp, so, sw);
    Graph g;
    g.n = n;
    g.nnz = sp[n];
    to_dev(g.indptr, sp.data(), n + 1, c.pool, c.stream);
    to_dev(g.indices, so.data(), g.nnz, c.pool, c.stream);
    to_dev(g.weights, sw.data(), g.nnz, c.pool, c.stream);
    Buf<float> de(c.pool, n * dim);
    if (init) EVOC_HIP(hipMemcpyAsync(de.p, init, n * dim * sizeof(float), hipMemcpyHostToDevice, c.stream));
    EmbedParams ep;
    ep.n_epochs = n_epochs;
    ep.noise_level = noise_level;
    ep.negative_sample_rate = negative_sample_rate;
    ep.initial_alpha = initial_alpha;
    ep.beta1 = beta1;
    ep.beta2 = beta2;
    ep.eps = eps;
    embed_device(g, dim, ep, de.p, init == nullptr, seed, c.stream, c.pool);
    EVOC_HIP(hipMemcpyAsync(out, de.p, n * dim * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    EVOC_HIP(hipStreamSynchronize(c.stream));
  });
}

int evoc_hip_label_prop_init(const int64_t* indptr, const int32_t* indices, const float* weights,
                             int64_t n, int32_t dim, const float* data, int64_t d, uint64_t seed,
                             float* out) {
  return guarded([&] {
    EVOC_CHECK(indptr && out, "label_prop_init: NULL buffer");
    EVOC_CHECK(n >= 1 && dim >= 1 && dim <= 16, "label_prop_init: dim must be in [1, 16]");
    check_finite_host(weights, indptr[n], "weights");
    if (data) check_finite_host(data, n * d, "data");
    Ctx& c = context(0);
    Buf<float> dx;
    if (data) to_dev(dx, data, n * d, c.pool, c.stream);
    label_prop_init_host(indptr, indices, weights, n, dim, data ? dx.p : nullptr, d, d, seed, out,
                         c.stream, c.pool, 0);
  });
}

int evoc_hip_core_sq(const float* emb, int64_t n, int32_t dim, int32_t min_samples,
                     float* core_sq) {
  return guarded([&] {
    EVOC_CHECK(emb && core_sq, "core_sq: NULL buffer");
    EVOC_CHECK(n >= 1 && dim >= 1 && dim <= 16, "core_sq: dim must be in [1, 16]");
    EVOC_CHECK(min_samples >= 0 && min_samples <= 64, "min_samples must be in [0, 64]");
    check_finite_host(emb, n * dim, "embedding");
    Ctx& c = context(0);
    Buf<float> de, dc(c.pool, n);
    to_dev(de, emb, n * dim, c.pool, c.stream);
    core_sq_device(de.p, n, dim, min_samples, dc.p, c.stream, c.pool);
    EVOC_HIP(hipMemcpyAsync(core_sq, dc.p, n * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    EVOC_HIP(hipStreamSynchronize(c.stream));
  });
}

int evoc_hip_mst(const float* emb, int64_t n, int32_t dim, int32_t min_samples, int32_t* u,
                 int32_t* v, float* w) {
  return guarded([&] {
    EVOC_CHECK(emb && u && v && w, "mst: NULL buffer");
    EVOC_CHECK(n >= 1 && dim >= 1 && dim <= 16, "mst: dim must be in [1, 16]");
    EVOC_CHECK(min_samples >= 0 && min_samples <= 64, "min_samples must be in [0, 64]");
    check_finite_host(emb, n * dim, "embedding");
    Ctx& c = context(0);
    Buf<float> de;
    to_dev(de, emb, n * dim, c.pool, c.stream);
    mst_device(de.p, n, dim, min_samples, u, v, w, c.stream, c.pool);
  });
}

int evoc_hip_condense_tree(const int64_t* left, const int64_t* right, const float* distance,
                           const int64_t* size, int64_t n, int64_t min_cluster_size,
                           int64_t* parent, int64_t* child, float* lambda, int64_t* child_size,
                           int64_t* n_rows) {
  return guarded([&] {
    EVOC_CHECK(n_rows && parent && child && lambda && child_size, "condense: NULL buffer");
    EVOC_CHECK(n >= 1 && n <= (int64_t(1) << 30), "condense: n must be in [1, 2^30]");
    EVOC_CHECK(n == 1 || (left && right && distance && size), "condense: NULL buffer");
    EVOC_CHECK(min_cluster_size >= 0, "condense: min_cluster_size must be >= 0");
    // the rows must form one binary tree over the n points: children of row i are
    // earlier nodes, each node is a child once, and sizes add up
    const int64_t nr = n - 1;
    std::vector<char> used(2 * n, 0);
    for (int64_t i = 0; i < nr; ++i) {
      const int64_t l = left[i], r = right[i];
      EVOC_CHECK(l >= 0 && r >= 0 && l < n + i && r < n + i && l != r && !used[l] && !used[r],
                 "condense: rows are not a binary tree in linkage order");
      used[l] = used[r] = 1;
      const int64_t ls = l < n ? 1 : size[l - n], rs = r < n ? 1 : size[r - n];
      EVOC_CHECK(size[i] == ls + rs, "condense: a row's size is not the sum of its children's");
      EVOC_CHECK(std::isfinite(distance[i]) && distance[i] >= 0.0f,
                 "condense: distances must be finite and >= 0");
    }
    Ctx& c = context(0);
    Linkage L;
    L.n = n;
    L.left.assign(left, left + nr);
    L.right.assign(right, right + nr);
    L.size.assign(size, size + nr);
    L.dist.assign(distance, distance + nr);
    const CTree t = condense_device(L, min_cluster_size, c.stream, c.pool);
    const size_t m = t.parent.size();
    std::memcpy(parent, t.parent.data(), m * sizeof(int64_t));
    std::memcpy(child, t.child.data(), m * sizeof(int64_t));
    std::memcpy(lambda, t.lambda.data(), m * sizeof(float));
    std::memcpy(child_size, t.size.data(), m * sizeof(int64_t));
    *n_rows = (int64_t)m;
  });
}

int evoc_hip_cluster_layers(const float* emb, int64_t n, int32_t dim, int32_t min_samples,
                            int64_t base_mcs, double min_sim, int64_t max_layers, int64_t* labels,
                            float* strengths, double* persistence, int64_t* n_layers) {
  return guarded([&] {
    EVOC_CHECK(emb && labels && strengths && persistence && n_layers, "layers: NULL buffer");
    EVOC_CHECK(n >= 1 && dim >= 1 && dim <= 16, "layers: dim must be in [1, 16]");
    EVOC_CHECK(min_samples >= 0 && min_samples <= 64, "min_samples must be in [0, 64]");
    check_finite_host(emb, n * dim, "embedding");
    Ctx& c = context(0);
    Buf<float> de;
    to_dev(de, emb, n * dim, c.pool, c.stream);
    evoc_params_t p;
    evoc_hip_default_params(&p);
    p.min_samples = min_samples;
    p.base_min_cluster_size = (int32_t)base_mcs;
    p.min_similarity_threshold = min_sim;
    p.max_layers = (int32_t)max_layers;
    Layers lay = cluster_device(de.p, n, dim, p, c.stream, c.pool, nullptr, nullptr);
    const int64_t nl = std::min<int64_t>(lay.nl(), std::max<int64_t>(max_layers, 1));
    std::memcpy(labels, lay.labels, (size_t)nl * n * sizeof(int64_t));
    std::memcpy(strengths, lay.strengths, (size_t)nl * n * sizeof(float));
    for (int64_t l = 0; l < nl; ++l) persistence[l] = lay.persistence[l];
    *n_layers = nl;
  });
}

int evoc_hip_search_n_clusters(const float* emb, int64_t n, int32_t dim, int32_t min_samples,
                               int64_t target_k, int64_t* labels, float* strengths) {
  return guarded([&] {
    EVOC_CHECK(emb && labels && strengths, "search: NULL buffer");
    EVOC_CHECK(n >= 1 && dim >= 1 && dim <= 16, "search: dim must be in [1, 16]");
    EVOC_CHECK(min_samples >= 0 && min_samples <= 64, "min_samples must be in [0, 64]");
    check_finite_host(emb, n * dim, "embedding");
    Ctx& c = context(0);
    Buf<float> de;
    to_dev(de, emb, n * dim, c.pool, c.stream);
    evoc_params_t p;
    evoc_hip_default_params(&p);
    p.min_samples = min_samples;
    p.approx_n_clusters = (int32_t)std::max<int64_t>(target_k, 1);
    Layers lay = cluster_device(de.p, n, dim, p, c.stream, c.pool, nullptr, nullptr);
    std::memcpy(labels, lay.labels, n * sizeof(int64_t));
    std::memcpy(strengths, lay.strengths, n * sizeof(float));
  });
}

int evoc_hip_mlp_loss_grads(const float* x, const float* t, int32_t rows, int32_t d,
                            int32_t out, const float* w0, const float* b0, const float* w1,
                            const float* b1, const float* w2, const float* b2, float* gw0,
                            float* gb0, float* gw1, float* gb1, float* gw2, float* gb2,
                            float* y, double* loss) {
  return guarded([&] {
    EVOC_CHECK(x && t && w0 && b0 && w1 && b1 && w2 && b2 && gw0 && gb0 && gw1 && gb1 && gw2 &&
                   gb2 && y && loss, "mlp: NULL buffer");
    EVOC_CHECK(rows >= 1 && d >= 1 && out >= 1 && out <= 32, "mlp: bad shape");
    Ctx& c = context(0);
    const float* w[3] = {w0, w1, w2};
    const float* b[3] = {b0, b1, b2};
    float* gw[3] = {gw0, gw1, gw2};
    float* gb[3] = {gb0, gb1, gb2};
    mlp_loss_and_grads(x, t, rows, d, out, w, b, gw, gb, y, loss, c.stream, c.pool);
  });
}

}  // extern "C"
