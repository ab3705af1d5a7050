// evoc oracle — CPU restatement of evoc-rs 0.2.7 (TEST INFRASTRUCTURE ONLY).
// See evoc_oracle.h for the canonicalisations. Compile with -ffp-contract=off so
// that every f32 expression rounds exactly as written (Rust does not contract).
#include "evoc_oracle.h"

#include <algorithm>
#include <atomic>
#include <memory>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <set>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- random numbers
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
void philox4x32_10(uint32_t ctr[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * ctr[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ ctr[1] ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ ctr[3] ^ k1;
    uint32_t n3 = (uint32_t)p0;
    ctr[0] = n0; ctr[1] = n1; ctr[2] = n2; ctr[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// stream tags (ASCII) used as the fourth counter word
constexpr uint32_t TAG_NEG = 0x454E4547u;   // embedding negatives
constexpr uint32_t TAG_EINI = 0x45494E49u;  // embedding random init
constexpr uint32_t TAG_LPSD = 0x4C505344u;  // label-prop seeds
constexpr uint32_t TAG_LPOL = 0x4C504F4Cu;  // label-prop outliers
constexpr uint32_t TAG_RINI = 0x52494E49u;  // label-prop random init

inline void philox(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                   uint32_t out[4]) {
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
  philox4x32_10(out, (uint32_t)seed, (uint32_t)(seed >> 32));
}
inline uint32_t rng_range(uint32_t r, uint64_t n) { return (uint32_t)(((uint64_t)r * n) >> 32); }
inline double rng_unit(const uint32_t r[4]) {  // [0,1) with 53 bits
  return (double)(((uint64_t)(r[0] >> 5) << 26) | (uint64_t)(r[1] >> 6)) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------- distances
// High-dimensional f32 dot / squared distance: 64 strided partial sums, then a
// halving tree (s = 32..1: acc[j] += acc[j+s]).
// (Full blocks of 64 first, the ragged tail after: the same order, and the lane loop of a
// full block vectorises.)
float canon_dot(const float* a, const float* b, int64_t d) {
  float acc[64];
  for (int j = 0; j < 64; ++j) acc[j] = 0.0f;
  const int64_t full = d / 64 * 64;
  for (int64_t t = 0; t < full; t += 64)
    for (int j = 0; j < 64; ++j) {
      float p = a[t + j] * b[t + j];
      acc[j] = acc[j] + p;
    }
  for (int j = 0; full + j < d; ++j) {
    float p = a[full + j] * b[full + j];
    acc[j] = acc[j] + p;
  }
  for (int s = 32; s >= 1; s >>= 1)
    for (int j = 0; j < s; ++j) acc[j] = acc[j] + acc[j + s];
  return acc[0];
}
float canon_sq(const float* a, const float* b, int64_t d) {
  float acc[64];
  for (int j = 0; j < 64; ++j) acc[j] = 0.0f;
  const int64_t full = d / 64 * 64;
  for (int64_t t = 0; t < full; t += 64)
    for (int j = 0; j < 64; ++j) {
      float df = a[t + j] - b[t + j];
      float p = df * df;
      acc[j] = acc[j] + p;
    }
  for (int j = 0; full + j < d; ++j) {
    float df = a[full + j] - b[full + j];
    float p = df * df;
    acc[j] = acc[j] + p;
  }
  for (int s = 32; s >= 1; s >>= 1)
    for (int j = 0; j < s; ++j) acc[j] = acc[j] + acc[j + s];
  return acc[0];
}
inline float cos_dist(float dot, float na, float nb) {
  float den = na * nb;
  if (!(den > 0.0f)) return 1.0f;
  float c = dot / den;
  return 1.0f - c;
}

// Canonical f32 exponential of the fuzzy set: exp in f64, rounded once to f32. The
// reference calls f32::exp (libm expf, within ~0.5 ulp of this); the correctly rounded
// value is the platform-independent choice both sides share.
inline float cexp(float x) { return (float)std::exp((double)x); }
// fp64 instantiations of the fuzzy set (error statements against fp64) use exp itself
inline double cexp(double x) { return std::exp(x); }

// Low-dimensional squared distance (embedding space): sequential sum.
template <typename T>
inline T dsq_seq(const T* a, const T* b, int dim) {
  T s = T(0);
  for (int i = 0; i < dim; ++i) {
    T df = a[i] - b[i];
    T p = df * df;
    s = s + p;
  }
  return s;
}

}  // namespace

extern "C" uint32_t orc_philox_u32(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2,
                                   uint32_t c3) {
  uint32_t r[4];
  philox(seed, c0, c1, c2, c3, r);
  return r[0];
}

extern "C" int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

extern "C" float orc_canon_dist(const float* a, const float* b, int64_t d, int metric) {
  if (metric == 1) {
    float na = std::sqrt(canon_dot(a, a, d));
    float nb = std::sqrt(canon_dot(b, b, d));
    return cos_dist(canon_dot(a, b, d), na, nb);
  }
  return std::sqrt(canon_sq(a, b, d));
}

// ================================================================ A1: exhaustive kNN
// Self-join top-k by (dist, idx) with self removed by j != i
// (src/nearest_neighbours/nearest_neighbour_gpu.rs:291-303; CPU drop-self at
// src/nearest_neighbours/nearest_neighbour_cpu.rs:350-361).
extern "C" int orc_knn_exhaustive(const float* x, int64_t n, int64_t d, int metric, int k,
                                  int32_t* idx_out, float* dist_out) {
  if (n <= 0 || k <= 0) return 0;
  std::vector<float> nrm;
  if (metric == 1) {
    nrm.resize(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) nrm[i] = std::sqrt(canon_dot(x + i * d, x + i * d, d));
  }
  const int kk = (int)std::min<int64_t>(k, n - 1);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < n; ++i) {
    std::vector<std::pair<float, int32_t>> best;  // sorted ascending (dist, idx)
    best.reserve(kk + 1);
    const float* xi = x + i * d;
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      float dist = metric == 1 ? cos_dist(canon_dot(xi, x + j * d, d), nrm[i], nrm[j])
                               : std::sqrt(canon_sq(xi, x + j * d, d));
      std::pair<float, int32_t> c(dist, (int32_t)j);
      if ((int)best.size() == kk && !(c < best.back())) continue;
      auto it = std::upper_bound(best.begin(), best.end(), c);
      best.insert(it, c);
      if ((int)best.size() > kk) best.pop_back();
    }
    for (int t = 0; t < k; ++t) {
      idx_out[i * k + t] = t < kk ? best[t].second : -1;
      dist_out[i * k + t] = t < kk ? best[t].first : INFINITY;
    }
  }
  return 0;
}

// Exhaustive kNN of selected query rows against the whole matrix (same canonical
// distance and (dist, idx) order as orc_knn_exhaustive); used to time the CPU on a
// bounded sample of rows.
extern "C" int orc_knn_rows(const float* x, int64_t n, int64_t d, int metric, int k,
                            const int64_t* rows, int64_t nrows, int32_t* idx_out,
                            float* dist_out) {
  std::vector<float> nrm;
  if (metric == 1) {
    nrm.resize(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) nrm[i] = std::sqrt(canon_dot(x + i * d, x + i * d, d));
  }
  const int kk = (int)std::min<int64_t>(k, n - 1);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t r = 0; r < nrows; ++r) {
    const int64_t i = rows[r];
    std::vector<std::pair<float, int32_t>> best;
    const float* xi = x + i * d;
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      float dist = metric == 1 ? cos_dist(canon_dot(xi, x + j * d, d), nrm[i], nrm[j])
                               : std::sqrt(canon_sq(xi, x + j * d, d));
      std::pair<float, int32_t> c(dist, (int32_t)j);
      if ((int)best.size() == kk && !(c < best.back())) continue;
      best.insert(std::upper_bound(best.begin(), best.end(), c), c);
      if ((int)best.size() > kk) best.pop_back();
    }
    for (int t = 0; t < k; ++t) {
      idx_out[r * k + t] = t < kk ? best[t].second : -1;
      dist_out[r * k + t] = t < kk ? best[t].first : INFINITY;
    }
  }
  return 0;
}

// ---------------------------------------------------------------- NNDescent
// Textbook NN-Descent (Dong, Charikar and Li, WWW 2011): random initial lists, then local
// joins of each node's sampled new and old neighbours (forward and reverse, rho * K of
// each) until fewer than delta * n * K list entries change. A stand-in for the reference's
// default backend (ann-search-rs nndescent, not vendored), used only to time the CPU
// baseline. Distances are the oracle's canonical ones, so a found neighbour carries the
// exact search's distance bits.
namespace {
struct NndEntry { float d; int32_t j; uint8_t fresh; };
struct SpinLock {
  std::atomic<int> f{0};
  void lock() { while (f.exchange(1, std::memory_order_acquire)) {} }
  void unlock() { f.store(0, std::memory_order_release); }
};
}  // namespace

extern "C" int orc_knn_nndescent(const float* x, int64_t n, int64_t d, int metric, int k,
                                 uint64_t seed, int max_iters, float delta, float rho,
                                 int32_t* idx_out, float* dist_out, int* iters_done) {
  const int K = (int)std::min<int64_t>(k, n - 1);
  if (K <= 0) {
    for (int64_t q = 0; q < n * k; ++q) { idx_out[q] = -1; dist_out[q] = INFINITY; }
    if (iters_done) *iters_done = 0;
    return 0;
  }
  std::vector<float> nrm(n, 0.0f);
  if (metric == 1) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) nrm[i] = std::sqrt(canon_dot(x + i * d, x + i * d, d));
  }
  auto dist = [&](int64_t a, int64_t b) {
    return metric == 1 ? cos_dist(canon_dot(x + a * d, x + b * d, d), nrm[a], nrm[b])
                       : std::sqrt(canon_sq(x + a * d, x + b * d, d));
  };
  std::vector<NndEntry> L(n * K);
  std::unique_ptr<SpinLock[]> locks(new SpinLock[n]);
  // insert (dd, j) into i's sorted list unless present or not better than the last
  auto push = [&](int64_t i, float dd, int32_t j) -> int {
    NndEntry* r = &L[i * K];
    if (!(dd < r[K - 1].d || (dd == r[K - 1].d && j < r[K - 1].j))) return 0;
    locks[i].lock();
    int res = 0;
    if (dd < r[K - 1].d || (dd == r[K - 1].d && j < r[K - 1].j)) {
      bool dup = false;
      for (int q = 0; q < K; ++q) dup |= r[q].j == j;
      if (!dup) {
        int q = K - 1;
        while (q > 0 && (r[q - 1].d > dd || (r[q - 1].d == dd && r[q - 1].j > j))) {
          r[q] = r[q - 1];
          --q;
        }
        r[q] = {dd, j, 1};
        res = 1;
      }
    }
    locks[i].unlock();
    return res;
  };
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    for (int q = 0; q < K; ++q) L[i * K + q] = {INFINITY, INT32_MAX, 0};
    int got = 0;
    for (uint32_t t = 0; got < K; ++t) {
      uint32_t r[4];
      philox(seed, (uint32_t)i, (uint32_t)(i >> 32), t, 0x4E4E4449u, r);
      const int64_t j = rng_range(r[0], (uint64_t)n);
      if (j == i) continue;
      bool dup = false;
      for (int q = 0; q < K; ++q) dup |= L[i * K + q].j == (int32_t)j;
      if (dup) continue;
      NndEntry* row = &L[i * K];
      const float dd = dist(i, j);
      int q = K - 1;
      while (q > 0 && (row[q - 1].d > dd || (row[q - 1].d == dd && row[q - 1].j > (int32_t)j))) {
        row[q] = row[q - 1];
        --q;
      }
      row[q] = {dd, (int32_t)j, 1};
      ++got;
    }
  }
  const int S = std::max(1, (int)(rho * (float)K));
  std::vector<int32_t> nf(n * S), of(n * K);
  std::vector<int32_t> nfc(n), ofc(n);
  std::vector<int32_t> nr(n * S), orv(n * S);
  std::vector<std::atomic<int32_t>> nrc(n), orc(n);
  int it = 0;
  for (; it < max_iters; ++it) {
    // forward samples: up to S new entries per node (marked old), all old entries
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      int a = 0, b = 0;
      for (int q = 0; q < K; ++q) {
        NndEntry& e = L[i * K + q];
        if (e.j == INT32_MAX) continue;
        if (e.fresh && a < S) {
          nf[i * S + a++] = e.j;
          e.fresh = 0;
        } else if (!e.fresh) {
          of[i * K + b++] = e.j;
        }
      }
      nfc[i] = a;
      ofc[i] = b;
      nrc[i].store(0, std::memory_order_relaxed);
      orc[i].store(0, std::memory_order_relaxed);
    }
    // reverse samples (the first S arrivals of each kind)
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      for (int a = 0; a < nfc[i]; ++a) {
        const int32_t j = nf[i * S + a];
        const int32_t pos = nrc[j].fetch_add(1, std::memory_order_relaxed);
        if (pos < S) nr[(int64_t)j * S + pos] = (int32_t)i;
      }
      for (int b = 0; b < std::min(ofc[i], S); ++b) {
        const int32_t j = of[i * K + b];
        const int32_t pos = orc[j].fetch_add(1, std::memory_order_relaxed);
        if (pos < S) orv[(int64_t)j * S + pos] = (int32_t)i;
      }
    }
    int64_t updates = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : updates)
    for (int64_t i = 0; i < n; ++i) {
      std::vector<int32_t> nw, od;
      for (int a = 0; a < nfc[i]; ++a) nw.push_back(nf[i * S + a]);
      for (int a = 0; a < std::min<int32_t>(nrc[i].load(), S); ++a) nw.push_back(nr[i * S + a]);
      for (int b = 0; b < ofc[i]; ++b) od.push_back(of[i * K + b]);
      for (int b = 0; b < std::min<int32_t>(orc[i].load(), S); ++b) od.push_back(orv[i * S + b]);
      std::sort(nw.begin(), nw.end());
      nw.erase(std::unique(nw.begin(), nw.end()), nw.end());
      std::sort(od.begin(), od.end());
      od.erase(std::unique(od.begin(), od.end()), od.end());
      for (size_t a = 0; a < nw.size(); ++a) {
        for (size_t b = a + 1; b < nw.size(); ++b) {
          const float dd = dist(nw[a], nw[b]);
          updates += push(nw[a], dd, nw[b]) + push(nw[b], dd, nw[a]);
        }
        for (size_t b = 0; b < od.size(); ++b) {
          if (od[b] == nw[a]) continue;
          const float dd = dist(nw[a], od[b]);
          updates += push(nw[a], dd, od[b]) + push(od[b], dd, nw[a]);
        }
      }
    }
    if ((double)updates < (double)delta * (double)n * (double)K) {
      ++it;
      break;
    }
  }
  for (int64_t i = 0; i < n; ++i)
    for (int q = 0; q < k; ++q) {
      const bool ok = q < K && L[i * K + q].j != INT32_MAX;
      idx_out[i * k + q] = ok ? L[i * K + q].j : -1;
      dist_out[i * k + q] = ok ? L[i * K + q].d : INFINITY;
    }
  if (iters_done) *iters_done = it;
  return 0;
}

// ================================================================ A2-A5: fuzzy set
// smooth_knn_dist: src/graph/fuzzy_graph.rs:33-98 (local_connectivity = 1, n_iter = 64).
// T = float is the reference's arithmetic; T = double is the fp64 restatement the error
// statements compare against.
template <typename T>
static void smooth_knn_dist(const T* dists, int64_t n, int k, size_t k_target, T* sigma,
                            T* rho) {
  const double target = std::log2((double)k_target);
  const double tol = 1e-5;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const T* d = dists + i * k;
    T r = d[0];  // local_connectivity = 1: index 0, no interpolation
    T lo = T(0), hi = std::numeric_limits<T>::max(), mid = T(1);
    for (int it = 0; it < 64; ++it) {
      T val = T(0);
      for (int t = 0; t < k; ++t) {
        T adjusted = std::fmax(d[t] - r, T(0));
        val += cexp(-(adjusted / mid));
      }
      if (std::fabs((double)val - target) < tol) break;
      if ((double)val > target) {
        hi = mid;
        mid = (lo + hi) / T(2);
      } else {
        lo = mid;
        if (hi == std::numeric_limits<T>::max()) mid = mid * T(2);
        else mid = (lo + hi) / T(2);
      }
    }
    sigma[i] = mid;
    rho[i] = r;
  }
}

template <typename T>
static int64_t symmetrise_coo(const std::vector<int64_t>& rr, const std::vector<int32_t>& cc,
                              const std::vector<T>& ww, int64_t n, int64_t* indptr,
                              int32_t* indices, T* weights);

// build_fuzzy_simplicial_set: src/graph/fuzzy_graph.rs:275-295, knn_to_coo :114-159,
// symmetrise_graph :174-228 (rows emitted sorted by column).
template <typename T>
static int64_t fuzzy_graph_impl(const int32_t* knn_idx, const T* knn_dist, int64_t n, int k,
                                float effective_k, int symmetrise, T* sigma_out, T* rho_out,
                                int64_t* indptr, int32_t* indices, T* weights) {
  // effective_n_neighbors.to_usize().unwrap_or(k): truncation when > -1
  size_t kt = (effective_k > -1.0f && std::isfinite(effective_k)) ? (size_t)effective_k : (size_t)k;
  std::vector<T> sigma(n), rho(n);
  smooth_knn_dist(knn_dist, n, k, kt, sigma.data(), rho.data());
  if (sigma_out) std::memcpy(sigma_out, sigma.data(), n * sizeof(T));
  if (rho_out) std::memcpy(rho_out, rho.data(), n * sizeof(T));
  // knn_to_coo
  std::vector<int64_t> rr;
  std::vector<int32_t> cc;
  std::vector<T> ww;
  for (int64_t i = 0; i < n; ++i)
    for (int t = 0; t < k; ++t) {
      int64_t j = knn_idx[i * k + t];
      if (j == i || j < 0) continue;
      T adjusted = std::fmax(knn_dist[i * k + t] - rho[i], T(0));
      T w;
      if (sigma[i] > T(0)) w = cexp(-(adjusted / sigma[i]));
      else if (adjusted > T(0)) w = T(0);
      else w = T(1);
      rr.push_back(i);
      cc.push_back((int32_t)j);
      ww.push_back(w);
    }
  if (!symmetrise) {
    indptr[0] = 0;
    int64_t p = 0;
    for (int64_t i = 0, e = 0; i < n; ++i) {
      while (e < (int64_t)rr.size() && rr[e] == i) {
        indices[p] = cc[e];
        weights[p] = ww[e];
        ++p; ++e;
      }
      indptr[i + 1] = p;
    }
    return p;
  }
  return symmetrise_coo(rr, cc, ww, n, indptr, indices, weights);
}

extern "C" int64_t orc_fuzzy_graph(const int32_t* knn_idx, const float* knn_dist, int64_t n,
                                   int k, float effective_k, int symmetrise, float* sigma_out,
                                   float* rho_out, int64_t* indptr, int32_t* indices,
                                   float* weights) {
  return fuzzy_graph_impl<float>(knn_idx, knn_dist, n, k, effective_k, symmetrise, sigma_out,
                                 rho_out, indptr, indices, weights);
}

// the same steps in fp64 (distances given in fp64)
extern "C" int64_t orc_fuzzy_graph_f64(const int32_t* knn_idx, const double* knn_dist,
                                       int64_t n, int k, float effective_k, int symmetrise,
                                       double* sigma_out, double* rho_out, int64_t* indptr,
                                       int32_t* indices, double* weights) {
  return fuzzy_graph_impl<double>(knn_idx, knn_dist, n, k, effective_k, symmetrise, sigma_out,
                                  rho_out, indptr, indices, weights);
}

// symmetrise_graph (src/graph/fuzzy_graph.rs:174-228) of a directed COO list: fuzzy union
// w_ij + w_ji - w_ij * w_ji, entries <= 0 dropped, both directions stored, rows sorted by
// column. A repeated (i, j) keeps its last weight.
template <typename T>
static int64_t symmetrise_coo(const std::vector<int64_t>& rr, const std::vector<int32_t>& cc,
                              const std::vector<T>& ww, int64_t n, int64_t* indptr,
                              int32_t* indices, T* weights) {
  // Every directed entry (i -> j, w) is filed under row i as forward and under row j as
  // backward; records keep their input order within a row (counting sort), then each row is
  // sorted stably by column, so for a repeated pair the last record of a direction wins.
  struct Rec { int32_t col; int32_t dir; T w; };
  const int64_t m = (int64_t)rr.size();
  std::vector<int64_t> off(n + 1, 0);
  for (int64_t e = 0; e < m; ++e) {
    off[rr[e] + 1]++;
    off[cc[e] + 1]++;
  }
  for (int64_t i = 0; i < n; ++i) off[i + 1] += off[i];
  std::vector<Rec> rec(2 * m);
  std::vector<int64_t> cur(off.begin(), off.end() - 1);
  for (int64_t e = 0; e < m; ++e) {
    rec[cur[rr[e]]++] = {cc[e], 0, ww[e]};
    rec[cur[cc[e]]++] = {(int32_t)rr[e], 1, ww[e]};
  }
  std::vector<int64_t> cnt(n, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; ++i) {
    std::stable_sort(rec.begin() + off[i], rec.begin() + off[i + 1],
                     [](const Rec& a, const Rec& b) { return a.col < b.col; });
    // compact in place: one record per column with (w_ij, w_ji) folded
    int64_t out = off[i];
    for (int64_t q = off[i]; q < off[i + 1];) {
      const int32_t j = rec[q].col;
      T wij = T(0), wji = T(0);
      for (; q < off[i + 1] && rec[q].col == j; ++q) {
        if (rec[q].dir == 0) wij = rec[q].w;
        else wji = rec[q].w;
      }
      T s = wij + wji;
      T pr = wij * wji;
      T wsym = s - pr;
      if (wsym > T(0)) rec[out++] = {j, 0, wsym};
    }
    cnt[i] = out - off[i];
  }
  indptr[0] = 0;
  for (int64_t i = 0; i < n; ++i) indptr[i + 1] = indptr[i] + cnt[i];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    for (int64_t q = 0; q < cnt[i]; ++q) {
      indices[indptr[i] + q] = rec[off[i] + q].col;
      weights[indptr[i] + q] = rec[off[i] + q].w;
    }
  return indptr[n];
}

// symmetrise_coo on explicit directed weights (for the reference's union tests)
extern "C" int64_t orc_symmetrise(const int64_t* rows, const int32_t* cols, const float* vals,
                                  int64_t nnz, int64_t n, int64_t* indptr, int32_t* indices,
                                  float* weights) {
  std::vector<int64_t> rr(rows, rows + nnz);
  std::vector<int32_t> cc(cols, cols + nnz);
  std::vector<float> ww(vals, vals + nnz);
  return symmetrise_coo(rr, cc, ww, n, indptr, indices, weights);
}

// ================================================================ A7: node embedding
// evoc_embedding: src/graph/embedding.rs:48-298; negatives drawn from Philox keyed by
// (seed, node, epoch, draw) instead of a per-node SmallRng stream.
static void embedding_impl(const std::vector<std::vector<std::pair<int32_t, float>>>& graph,
                           int dim, int n_epochs, float noise, float neg_rate, float alpha,
                           float beta1, float beta2, float eps, const float* init,
                           uint64_t seed, float* out) {
  const int64_t n = (int64_t)graph.size();
  if (n == 0) return;
  std::vector<float> embd(n * dim);
  if (init) {
    std::memcpy(embd.data(), init, n * dim * sizeof(float));
  } else {
    for (int64_t i = 0; i < n; ++i)
      for (int c = 0; c < dim; ++c) {
        uint32_t r[4];
        philox(seed, (uint32_t)i, (uint32_t)c, 0, TAG_EINI, r);
        embd[i * dim + c] = (float)(-0.25 + 0.5 * rng_unit(r));
      }
  }
  struct Edge { int32_t i, j; float w; };
  std::vector<Edge> edges;
  std::vector<int64_t> degree(n, 0);
  for (int64_t i = 0; i < n; ++i)
    for (auto& jw : graph[i])
      if (i < jw.first) {
        edges.push_back({(int32_t)i, jw.first, jw.second});
        degree[i]++;
        degree[jw.first]++;
      }
  if (edges.empty()) {
    std::memcpy(out, embd.data(), n * dim * sizeof(float));
    return;
  }
  const int64_t ne = (int64_t)edges.size();
  float max_w = 0.0f;
  for (auto& e : edges) max_w = std::fmax(max_w, e.w);
  std::vector<float> eps_s(ne), next_s(ne), eps_n(ne), next_n(ne);
  const float neg_factor = 1.5f / neg_rate;
  for (int64_t e = 0; e < ne; ++e) {
    float norm = edges[e].w / max_w;
    eps_s[e] = norm > 0.0f ? 1.0f / norm : 1e8f;
    next_s[e] = eps_s[e];
    eps_n[e] = eps_s[e] * neg_factor;
    next_n[e] = eps_n[e];
  }
  std::vector<int64_t> off(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) off[i + 1] = off[i] + degree[i];
  struct Slot { int64_t e; bool smaller; int32_t other; };
  std::vector<Slot> csr(2 * ne);
  std::vector<int64_t> cur(off.begin(), off.end() - 1);
  for (int64_t e = 0; e < ne; ++e) {
    csr[cur[edges[e].i]++] = {e, true, edges[e].j};
    csr[cur[edges[e].j]++] = {e, false, edges[e].i};
  }
  std::vector<float> m(n * dim, 0.0f), v(n * dim, 0.0f), grads(n * dim, 0.0f);
  std::vector<char> active(n, 0);
  const float one_m_b1 = 1.0f - beta1, one_m_b2 = 1.0f - beta2;
  const float denom = (float)std::max(n_epochs - 1, 1);
  const float n_epochs_f = (float)n_epochs;
  for (int epoch = 0; epoch < n_epochs; ++epoch) {
    const float lr = alpha * (1.0f - (float)epoch / n_epochs_f);
    const float epoch_t = (float)epoch;
    const float t = (float)(epoch + 1);
    const float b1t = std::pow(beta1, t), b2t = std::pow(beta2, t);
    const float sqrt_b2 = std::sqrt(1.0f - b2t);
    const float ad_scale = sqrt_b2 / (1.0f - b1t);
    const float epsc = sqrt_b2 * eps;
    const float gamma = 0.5f + (float)epoch / denom;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ni = 0; ni < n; ++ni) {
      float* g = &grads[ni * dim];
      for (int c = 0; c < dim; ++c) g[c] = 0.0f;
      const float* ei = &embd[ni * dim];
      bool touched = false;
      uint32_t draw = 0;
      for (int64_t s = off[ni]; s < off[ni + 1]; ++s) {
        const Slot& sl = csr[s];
        if (next_s[sl.e] > epoch_t) continue;
        touched = true;
        const float* eo = &embd[(int64_t)sl.other * dim];
        float dsq = dsq_seq(ei, eo, dim);
        if (dsq >= 1e-8f) {
          float dist = std::sqrt(dsq);
          float num = -((2.0f * noise) * dist) - 2.0f;
          float den = ((2.0f * dsq) - (0.5f * dist)) + 1.0f;
          float coeff = num / den;
          for (int c = 0; c < dim; ++c) g[c] = g[c] + coeff * (ei[c] - eo[c]);
        }
        if (sl.smaller) {
          float f = std::floor((epoch_t - next_n[sl.e]) / eps_n[sl.e]);
          uint64_t n_neg = (f >= 0.0f && f < 1.8e19f) ? (uint64_t)f : 0;
          for (uint64_t q = 0; q < n_neg; ++q) {
            uint32_t r[4];
            philox(seed, draw++, (uint32_t)epoch, (uint32_t)ni, TAG_NEG, r);
            int64_t kk = rng_range(r[0], (uint64_t)n);
            if (kk == ni) continue;
            const float* ek = &embd[kk * dim];
            float dsq_k = dsq_seq(ei, ek, dim);
            if (dsq_k > 0.01f) {
              float gc = (gamma * 4.0f) / ((1.0f + 0.25f * dsq_k) * dsq_k);
              for (int c = 0; c < dim; ++c) {
                float diff = ei[c] - ek[c];
                g[c] = g[c] + std::fmin(std::fmax(gc * diff, -4.0f), 4.0f);
              }
            }
          }
        }
      }
      active[ni] = touched;
    }
#pragma omp parallel for schedule(static)
    for (int64_t ni = 0; ni < n; ++ni) {
      if (!active[ni]) continue;
      for (int c = 0; c < dim; ++c) {
        int64_t q = ni * dim + c;
        float gd = grads[q];
        m[q] = m[q] + one_m_b1 * (gd - m[q]);
        v[q] = v[q] + one_m_b2 * (gd * gd - v[q]);
        float step = ((lr * ad_scale) * m[q]) / (std::sqrt(v[q]) + epsc);
        embd[q] = embd[q] + step;
      }
    }
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
      if (next_s[e] <= epoch_t) {
        next_s[e] = next_s[e] + eps_s[e];
        float f = std::floor((epoch_t - next_n[e]) / eps_n[e]);
        uint64_t cnt = (f >= 0.0f && f < 1.8e19f) ? (uint64_t)f : 0;
        next_n[e] = next_n[e] + (float)cnt * eps_n[e];
      }
    }
  }
  std::memcpy(out, embd.data(), n * dim * sizeof(float));
}

static std::vector<std::vector<std::pair<int32_t, float>>> csr_to_adj(const int64_t* indptr,
                                                                      const int32_t* indices,
                                                                      const float* w,
                                                                      int64_t n) {
  std::vector<std::vector<std::pair<int32_t, float>>> adj(n);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t p = indptr[i]; p < indptr[i + 1]; ++p) adj[i].push_back({indices[p], w[p]});
  return adj;
}

extern "C" int orc_embedding(const int64_t* indptr, const int32_t* indices,
                             const float* weights, int64_t n, int dim, int n_epochs,
                             float noise_level, float negative_sample_rate,
                             float initial_alpha, float beta1, float beta2, float eps,
                             const float* init, uint64_t seed, float* out) {
  auto adj = csr_to_adj(indptr, indices, weights, n);
  embedding_impl(adj, dim, n_epochs, noise_level, negative_sample_rate, initial_alpha, beta1,
                 beta2, eps, init, seed, out);
  return 0;
}

// ================================================================ A6: label-prop init
// src/graph/label_prop.rs and the CSR algebra of src/utils/sparse.rs.
namespace {
struct Csr {
  std::vector<int64_t> indptr;
  std::vector<int64_t> indices;
  std::vector<float> data;
  int64_t nrows = 0, ncols = 0;
};

// Csr::from_partition (src/utils/sparse.rs:246-255)
Csr from_partition(const std::vector<int64_t>& part, int64_t n_parts) {
  Csr c;
  c.nrows = (int64_t)part.size();
  c.ncols = n_parts;
  c.indptr.resize(c.nrows + 1);
  for (int64_t i = 0; i <= c.nrows; ++i) c.indptr[i] = i;
  c.indices = part;
  c.data.assign(c.nrows, 1.0f);
  return c;
}
// Csr::transpose (src/utils/sparse.rs:267-302): counting sort, rows ascending
Csr transpose(const Csr& a) {
  Csr t;
  t.nrows = a.ncols;
  t.ncols = a.nrows;
  t.indptr.assign(a.ncols + 1, 0);
  for (auto c : a.indices) t.indptr[c + 1]++;
  for (int64_t i = 0; i < a.ncols; ++i) t.indptr[i + 1] += t.indptr[i];
  t.indices.resize(a.indices.size());
  t.data.resize(a.data.size());
  std::vector<int64_t> cur(t.indptr.begin(), t.indptr.end() - 1);
  for (int64_t r = 0; r < a.nrows; ++r)
    for (int64_t p = a.indptr[r]; p < a.indptr[r + 1]; ++p) {
      int64_t c = a.indices[p];
      int64_t pos = cur[c]++;
      t.data[pos] = a.data[p];
      t.indices[pos] = r;
    }
  return t;
}
// Csr::matmul (src/utils/sparse.rs:318-370): per-row dense accumulator, output sorted
Csr matmul(const Csr& a, const Csr& b) {
  Csr c;
  c.nrows = a.nrows;
  c.ncols = b.ncols;
  std::vector<std::vector<std::pair<int64_t, float>>> rows(a.nrows);
#pragma omp parallel
  {
    std::vector<float> vals(b.ncols, 0.0f);
    std::vector<char> flag(b.ncols, 0);
    std::vector<int64_t> touched;
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < a.nrows; ++i) {
      touched.clear();
      for (int64_t p = a.indptr[i]; p < a.indptr[i + 1]; ++p) {
        int64_t k = a.indices[p];
        float av = a.data[p];
        for (int64_t q = b.indptr[k]; q < b.indptr[k + 1]; ++q) {
          int64_t col = b.indices[q];
          float prod = av * b.data[q];
          if (!flag[col]) {
            flag[col] = 1;
            touched.push_back(col);
            vals[col] = prod;
          } else {
            vals[col] = vals[col] + prod;
          }
        }
      }
      std::sort(touched.begin(), touched.end());
      auto& row = rows[i];
      row.reserve(touched.size());
      for (auto col : touched) {
        row.push_back({col, vals[col]});
        flag[col] = 0;
        vals[col] = 0.0f;
      }
    }
  }
  c.indptr.assign(a.nrows + 1, 0);
  for (int64_t i = 0; i < a.nrows; ++i) {
    for (auto& cv : rows[i]) {
      c.indices.push_back(cv.first);
      c.data.push_back(cv.second);
    }
    c.indptr[i + 1] = (int64_t)c.data.size();
  }
  return c;
}
// Csr::elementwise_mul (src/utils/sparse.rs:383-427)
Csr hadamard(const Csr& a, const Csr& b) {
  Csr c;
  c.nrows = a.nrows;
  c.ncols = a.ncols;
  c.indptr.assign(a.nrows + 1, 0);
  for (int64_t i = 0; i < a.nrows; ++i) {
    int64_t p = a.indptr[i], ep = a.indptr[i + 1], q = b.indptr[i], eq = b.indptr[i + 1];
    while (p < ep && q < eq) {
      int64_t ci = a.indices[p], cj = b.indices[q];
      if (ci == cj) {
        c.indices.push_back(ci);
        c.data.push_back(a.data[p] * b.data[q]);
        ++p; ++q;
      } else if (ci < cj) ++p;
      else ++q;
    }
    c.indptr[i + 1] = (int64_t)c.data.size();
  }
  return c;
}
// Csr::normalise_cols_l2 (src/utils/sparse.rs:437-469)
Csr norm_cols_l2(const Csr& a) {
  std::vector<float> sq(a.ncols, 0.0f);
  for (size_t p = 0; p < a.data.size(); ++p) sq[a.indices[p]] += a.data[p] * a.data[p];
  std::vector<float> inv(a.ncols);
  for (int64_t c = 0; c < a.ncols; ++c) {
    float nrm = std::sqrt(sq[c]);
    inv[c] = nrm > 0.0f ? 1.0f / nrm : 1.0f;
  }
  Csr r = a;
  for (size_t p = 0; p < r.data.size(); ++p) r.data[p] = a.data[p] * inv[a.indices[p]];
  return r;
}
// Csr::normalise_rows_l1 (src/utils/sparse.rs:479-501)
Csr norm_rows_l1(const Csr& a) {
  Csr r = a;
  for (int64_t i = 0; i < a.nrows; ++i) {
    float nrm = 0.0f;
    for (int64_t p = a.indptr[i]; p < a.indptr[i + 1]; ++p) nrm += std::fabs(a.data[p]);
    if (nrm > 0.0f) {
      float inv = 1.0f / nrm;
      for (int64_t p = a.indptr[i]; p < a.indptr[i + 1]; ++p) r.data[p] = r.data[p] * inv;
    }
  }
  return r;
}
// Csr::matmul_dense (src/utils/sparse.rs:542-567), row-major dense
std::vector<float> matmul_dense(const Csr& a, const std::vector<float>& rhs, int64_t d) {
  std::vector<float> out(a.nrows * d, 0.0f);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < a.nrows; ++i) {
    float* row = &out[i * d];
    for (int64_t p = a.indptr[i]; p < a.indptr[i + 1]; ++p) {
      int64_t j = a.indices[p];
      float v = a.data[p];
      const float* r = &rhs[j * d];
      for (int64_t c = 0; c < d; ++c) row[c] = row[c] + v * r[c];
    }
  }
  return out;
}

// Csr::clip_values (src/utils/sparse.rs:502-510)
void clip_values(Csr& a, float lo, float hi) {
  for (auto& v : a.data) {
    if (v < lo) v = lo;
    else if (v > hi) v = hi;
  }
}
// Csr::from_coo (src/utils/sparse.rs:194-235): sort by (row, col), fold duplicates by
// summing in sorted order (a stable sort, so duplicates add in input order)
Csr from_coo(const int64_t* rows, const int64_t* cols, const float* vals, int64_t nnz, int64_t n) {
  Csr c;
  c.nrows = c.ncols = n;
  c.indptr.assign(n + 1, 0);
  std::vector<int64_t> ord(nnz);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
    return rows[a] != rows[b] ? rows[a] < rows[b] : cols[a] < cols[b];
  });
  int64_t pr = -1, pc = -1;
  for (int64_t q : ord) {
    if (rows[q] == pr && cols[q] == pc) {
      c.data.back() = c.data.back() + vals[q];
    } else {
      c.data.push_back(vals[q]);
      c.indices.push_back(cols[q]);
      c.indptr[rows[q] + 1]++;
      pr = rows[q];
      pc = cols[q];
    }
  }
  for (int64_t i = 0; i < n; ++i) c.indptr[i + 1] += c.indptr[i];
  return c;
}

// label_prop_iteration: src/graph/label_prop.rs:45-86
std::vector<int64_t> lp_iteration(const Csr& g, const std::vector<int64_t>& labels) {
  std::vector<int64_t> out(g.nrows);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < g.nrows; ++i) {
    if (labels[i] >= 0) {
      out[i] = labels[i];
      continue;
    }
    std::vector<std::pair<int64_t, float>> votes;
    for (int64_t p = g.indptr[i]; p < g.indptr[i + 1]; ++p) {
      int64_t l = labels[g.indices[p]];
      float w = g.data[p];
      bool found = false;
      for (auto& v : votes)
        if (v.first == l) {
          v.second += w;
          found = true;
          break;
        }
      if (!found) votes.push_back({l, w});
    }
    int64_t best = -1;
    float max_vote = 1.0f;
    for (auto& v : votes) {
      if (v.first < 0) continue;
      if (v.second > max_vote) {
        max_vote = v.second;
        best = v.first;
      } else if (v.second == max_vote) {
        best = v.first;
      }
    }
    out[i] = best;
  }
  return out;
}
// label_outliers: src/graph/label_prop.rs:105-140
void lp_outliers(const Csr& g, std::vector<int64_t>& labels, uint64_t seed) {
  int64_t max_label = -1;
  for (auto l : labels) max_label = std::max(max_label, l);
  if (max_label < 0) return;
  for (int64_t i = 0; i < g.nrows; ++i) {
    if (labels[i] >= 0) continue;
    std::vector<int64_t> queue{i};
    bool found = false;
    for (int it = 0; it < 64; ++it) {
      if (queue.empty()) break;
      int64_t node = queue.back();
      queue.pop_back();
      for (int64_t p = g.indptr[node]; p < g.indptr[node + 1]; ++p) {
        int64_t j = g.indices[p];
        if (labels[j] >= 0) {
          labels[i] = labels[j];
          found = true;
          break;
        }
        queue.push_back(j);
      }
      if (found) break;
    }
    if (!found) {
      uint32_t r[4];
      philox(seed, (uint32_t)i, (uint32_t)(i >> 32), 0, TAG_LPOL, r);
      labels[i] = (int64_t)rng_range(r[0], (uint64_t)max_label + 1);
    }
  }
}
// remap_labels: src/graph/label_prop.rs:155-177
int64_t lp_remap(std::vector<int64_t>& labels) {
  std::vector<int64_t> s(labels);
  std::sort(s.begin(), s.end());
  s.erase(std::unique(s.begin(), s.end()), s.end());
  s.erase(std::remove_if(s.begin(), s.end(), [](int64_t l) { return l < 0; }), s.end());
  std::unordered_map<int64_t, int64_t> mp;
  for (size_t i = 0; i < s.size(); ++i) mp[s[i]] = (int64_t)i;
  int64_t next = (int64_t)s.size();
  for (auto& l : labels) l = l < 0 ? next++ : mp[l];
  return next;
}
// label_prop_loop: src/graph/label_prop.rs:202-227
std::vector<int64_t> lp_loop(const Csr& g, int64_t approx_parts, int n_iter, uint64_t seed,
                             int64_t* n_parts) {
  const int64_t n = g.nrows;
  std::vector<int64_t> labels(n, -1);
  for (int64_t i = 0; i < approx_parts; ++i) {
    uint32_t r[4];
    philox(seed, (uint32_t)i, (uint32_t)(i >> 32), 0, TAG_LPSD, r);
    labels[rng_range(r[0], (uint64_t)n)] = i;
  }
  for (int it = 0; it < n_iter; ++it) labels = lp_iteration(g, labels);
  lp_outliers(g, labels, seed + 2000);
  *n_parts = lp_remap(labels);
  return labels;
}

// normalise_to_unit_range: src/graph/label_prop.rs:239-266
void unit_range(std::vector<float>& m, int64_t n, int64_t d) {
  for (int64_t j = 0; j < d; ++j) {
    float lo = INFINITY, hi = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
      lo = std::fmin(lo, m[i * d + j]);
      hi = std::fmax(hi, m[i * d + j]);
    }
    float range = hi - lo;
    if (range > 0.0f) {
      float mid = (hi + lo) / 2.0f, half = range / 2.0f;
      for (int64_t i = 0; i < n; ++i) m[i * d + j] = (m[i * d + j] - mid) / half;
    }
  }
}

// Symmetric eigen-decomposition by cyclic Jacobi (small n); eigenvalues descending.
void jacobi_eigen(std::vector<double>& a, int64_t n, std::vector<double>& evals,
                  std::vector<double>& evecs) {
  evecs.assign(n * n, 0.0);
  for (int64_t i = 0; i < n; ++i) evecs[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int64_t p = 0; p < n; ++p)
      for (int64_t q = p + 1; q < n; ++q) off += a[p * n + q] * a[p * n + q];
    if (off < 1e-30) break;
    for (int64_t p = 0; p < n; ++p)
      for (int64_t q = p + 1; q < n; ++q) {
        double apq = a[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        double theta = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int64_t k = 0; k < n; ++k) {
          double akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = c * akp - s * akq;
          a[k * n + q] = s * akp + c * akq;
        }
        for (int64_t k = 0; k < n; ++k) {
          double apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = c * apk - s * aqk;
          a[q * n + k] = s * apk + c * aqk;
        }
        for (int64_t k = 0; k < n; ++k) {
          double vkp = evecs[k * n + p], vkq = evecs[k * n + q];
          evecs[k * n + p] = c * vkp - s * vkq;
          evecs[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int64_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t x, int64_t y) { return a[x * n + x] > a[y * n + y]; });
  std::vector<double> vecs(n * n);
  evals.resize(n);
  for (int64_t c = 0; c < n; ++c) {
    evals[c] = a[order[c] * n + order[c]];
    for (int64_t r = 0; r < n; ++r) vecs[r * n + c] = evecs[r * n + order[c]];
  }
  evecs.swap(vecs);
}

// pca_init: src/graph/label_prop.rs:288-321. U*S of the centred f64 matrix from the
// eigen-decomposition of its Gram matrix; each component's sign is fixed so that its
// largest-magnitude entry is positive (faer's sign convention is not pinned).
std::vector<float> pca_init(const std::vector<float>& data, int64_t n, int64_t d, int dim) {
  const int64_t k = std::min<int64_t>(std::min<int64_t>(dim, n), d);
  std::vector<double> means(d, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < d; ++j) means[j] += (double)data[i * d + j];
  const double inv_n = 1.0 / (double)n;
  for (auto& m : means) m *= inv_n;
  std::vector<double> c(n * d);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < d; ++j) c[i * d + j] = (double)data[i * d + j] - means[j];
  std::vector<double> gram(n * n, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < n; ++j) {
      double s = 0.0;
      for (int64_t t = 0; t < d; ++t) s += c[i * d + t] * c[j * d + t];
      gram[i * n + j] = s;
    }
  std::vector<double> ev, vec;
  jacobi_eigen(gram, n, ev, vec);
  std::vector<float> out(n * dim, 0.0f);
  for (int64_t comp = 0; comp < k; ++comp) {
    double sv = std::sqrt(std::max(ev[comp], 0.0));
    int64_t arg = 0;
    for (int64_t i = 1; i < n; ++i)
      if (std::fabs(vec[i * n + comp]) > std::fabs(vec[arg * n + comp])) arg = i;
    double sign = vec[arg * n + comp] < 0 ? -1.0 : 1.0;
    for (int64_t i = 0; i < n; ++i) out[i * dim + comp] = (float)(sign * vec[i * n + comp] * sv);
  }
  // normalise the k real components to [-1, 1]; padding columns (k < dim) stay 0
  std::vector<float> sub(n * k);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < k; ++j) sub[i * k + j] = out[i * dim + j];
  unit_range(sub, n, k);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < k; ++j) out[i * dim + j] = sub[i * k + j];
  return out;
}
// random_init: src/graph/label_prop.rs:338-362
std::vector<float> random_init(int64_t n, int dim, uint64_t seed) {
  std::vector<float> out(n * dim);
  for (int64_t i = 0; i < n; ++i) {
    double nrm = 0.0;
    std::vector<double> row(dim);
    for (int j = 0; j < dim; ++j) {
      uint32_t r[4];
      philox(seed, (uint32_t)i, (uint32_t)j, 0, TAG_RINI, r);
      row[j] = (double)(float)(rng_unit(r) * 2.0 - 1.0);
      nrm += row[j] * row[j];
    }
    nrm = std::sqrt(nrm);
    for (int j = 0; j < dim; ++j) out[i * dim + j] = nrm > 0.0 ? (float)(row[j] / nrm) : (float)row[j];
  }
  return out;
}

// The expansion of label_prop_init_inner (src/graph/label_prop.rs:487-516): the mean of
// rows_l1((G .* G^T) R_norm) L and rows_l1(R_norm) L, columns centred, times scaling.
std::vector<float> expand_layout(const Csr& g, const Csr& nrm, const std::vector<float>& layout,
                                 int dim, double scaling) {
  const int64_t n = g.nrows;
  Csr gt = transpose(g);
  Csr sym = hadamard(g, gt);
  Csr expander = norm_rows_l1(matmul(sym, nrm));
  Csr nrm_l1 = norm_rows_l1(nrm);
  std::vector<float> expanded = matmul_dense(expander, layout, dim);
  std::vector<float> direct = matmul_dense(nrm_l1, layout, dim);
  std::vector<float> res(n * dim);
  for (int64_t q = 0; q < n * dim; ++q) res[q] = (expanded[q] + direct[q]) * 0.5f;
  const float inv_n = (float)(1.0 / (double)n);
  for (int j = 0; j < dim; ++j) {
    float sum = 0.0f;
    for (int64_t i = 0; i < n; ++i) sum += res[i * dim + j];
    float mean = sum * inv_n;
    for (int64_t i = 0; i < n; ++i) res[i * dim + j] = res[i * dim + j] - mean;
  }
  const float sc = (float)scaling;
  for (auto& v : res) v = sc * v;
  return res;
}

// label_prop_init_inner: src/graph/label_prop.rs:403-517
std::vector<float> lp_init_inner(const Csr& g, int dim, const std::vector<float>* data,
                                 int64_t d, int n_emb_epochs, int64_t approx_parts, int n_lp_iter,
                                 int64_t base_thresh, double scaling, uint64_t seed) {
  const int64_t n = g.nrows;
  if (n < base_thresh) {
    if (data && (int64_t)data->size() == n * d && dim <= d) return pca_init(*data, n, d, dim);
    return random_init(n, dim, seed);
  }
  int64_t n_parts = 0;
  std::vector<int64_t> part = lp_loop(g, approx_parts, n_lp_iter, seed, &n_parts);
  Csr base = from_partition(part, n_parts);
  Csr nrm = norm_cols_l2(base);
  Csr nrt = transpose(nrm);
  Csr reducer = norm_rows_l1(nrt);
  Csr reduced = matmul(matmul(nrt, g), base);
  clip_values(reduced, 0.0f, 1.0f);
  std::vector<float> rdata;
  if (data) rdata = matmul_dense(reducer, *data, d);
  int ep = std::min(n_emb_epochs, 255);
  std::vector<float> rinit = lp_init_inner(reduced, dim, data ? &rdata : nullptr, d, ep,
                                           approx_parts / 4, n_lp_iter, base_thresh, scaling,
                                           seed + 3000);
  std::vector<std::vector<std::pair<int32_t, float>>> adj(reduced.nrows);
  for (int64_t i = 0; i < reduced.nrows; ++i)
    for (int64_t p = reduced.indptr[i]; p < reduced.indptr[i + 1]; ++p)
      adj[i].push_back({(int32_t)reduced.indices[p], reduced.data[p]});
  std::vector<float> layout(reduced.nrows * dim);
  embedding_impl(adj, dim, ep, 0.5f, 1.0f, (float)(0.001 * (double)n_emb_epochs), 0.5f, 0.9f,
                 1e-7f, rinit.data(), seed + 4000, layout.data());
  return expand_layout(g, nrm, layout, dim, scaling);
}
}  // namespace

// label_propagation_init: src/graph/label_prop.rs:537-564 (input graph given as CSR
// with rows sorted by column, i.e. Csr::from_coo of the fuzzy set)
extern "C" int orc_label_prop_init(const int64_t* indptr, const int32_t* indices,
                                   const float* weights, int64_t n, int dim, const float* data,
                                   int64_t d, uint64_t seed, float* out) {
  Csr g;
  g.nrows = g.ncols = n;
  g.indptr.assign(indptr, indptr + n + 1);
  g.indices.assign(indices, indices + indptr[n]);
  g.data.assign(weights, weights + indptr[n]);
  int64_t approx = (int64_t)(8.0 * std::sqrt((double)n));
  approx = std::min<int64_t>(std::max<int64_t>(approx, 256), 16384);
  approx = std::min<int64_t>(approx, n);
  std::vector<float> dv;
  if (data) dv.assign(data, data + n * d);
  std::vector<float> r =
      lp_init_inner(g, dim, data ? &dv : nullptr, d, 50, approx, 20, 64, 0.1, seed);
  std::memcpy(out, r.data(), n * dim * sizeof(float));
  return 0;
}

// ---------------------------------------------------------------- test entry points
// The label-prop pieces and the CSR algebra one by one, so that the reference's own unit
// tests (src/graph/label_prop.rs:574-738, src/utils/sparse.rs:656-951) pin them.
namespace {
Csr csr_in(const int64_t* ip, const int64_t* ix, const float* dv, int64_t nr, int64_t nc) {
  Csr c;
  c.nrows = nr;
  c.ncols = nc;
  c.indptr.assign(ip, ip + nr + 1);
  c.indices.assign(ix, ix + ip[nr]);
  c.data.assign(dv, dv + ip[nr]);
  return c;
}
int64_t csr_out(const Csr& c, int64_t* ip, int64_t* ix, float* dv, int64_t cap, int64_t* shape) {
  shape[0] = c.nrows;
  shape[1] = c.ncols;
  const int64_t nnz = (int64_t)c.data.size();
  if (nnz > cap) return -1;
  std::copy(c.indptr.begin(), c.indptr.end(), ip);
  std::copy(c.indices.begin(), c.indices.end(), ix);
  std::copy(c.data.begin(), c.data.end(), dv);
  return nnz;
}
}  // namespace

// op: 0 transpose(A), 1 A*B, 2 A .* B, 3 normalise_cols_l2(A), 4 normalise_rows_l1(A),
// 5 clip_values(A, 0, 1). Returns the stored count, or -1 if it exceeds cap.
extern "C" int64_t orc_csr_op(int op, const int64_t* aip, const int64_t* aix, const float* adv,
                              int64_t anr, int64_t anc, const int64_t* bip, const int64_t* bix,
                              const float* bdv, int64_t bnr, int64_t bnc, int64_t* oip,
                              int64_t* oix, float* odv, int64_t cap, int64_t* shape) {
  Csr a = csr_in(aip, aix, adv, anr, anc);
  Csr r;
  switch (op) {
    case 0: r = transpose(a); break;
    case 1: r = matmul(a, csr_in(bip, bix, bdv, bnr, bnc)); break;
    case 2: r = hadamard(a, csr_in(bip, bix, bdv, bnr, bnc)); break;
    case 3: r = norm_cols_l2(a); break;
    case 4: r = norm_rows_l1(a); break;
    case 5: r = a; clip_values(r, 0.0f, 1.0f); break;
    default: return -1;
  }
  return csr_out(r, oip, oix, odv, cap, shape);
}

extern "C" int64_t orc_csr_from_coo(const int64_t* rows, const int64_t* cols, const float* vals,
                                    int64_t nnz, int64_t n, int64_t* oip, int64_t* oix,
                                    float* odv) {
  int64_t shape[2];
  return csr_out(from_coo(rows, cols, vals, nnz, n), oip, oix, odv, nnz, shape);
}

extern "C" int orc_csr_from_partition(const int64_t* part, int64_t n, int64_t n_parts,
                                      int64_t* oip, int64_t* oix, float* odv) {
  int64_t shape[2];
  return csr_out(from_partition(std::vector<int64_t>(part, part + n), n_parts), oip, oix, odv, n,
                 shape) < 0 ? -1 : 0;
}

extern "C" int orc_csr_matmul_dense(const int64_t* ip, const int64_t* ix, const float* dv,
                                    int64_t nr, int64_t nc, const float* rhs, int64_t d,
                                    float* out) {
  std::vector<float> r(rhs, rhs + nc * d);
  std::vector<float> o = matmul_dense(csr_in(ip, ix, dv, nr, nc), r, d);
  std::copy(o.begin(), o.end(), out);
  return 0;
}

extern "C" int orc_lp_iteration(const int64_t* ip, const int64_t* ix, const float* dv, int64_t n,
                                const int64_t* labels, int64_t* out) {
  std::vector<int64_t> l(labels, labels + n);
  std::vector<int64_t> r = lp_iteration(csr_in(ip, ix, dv, n, n), l);
  std::copy(r.begin(), r.end(), out);
  return 0;
}

extern "C" int orc_lp_outliers(const int64_t* ip, const int64_t* ix, const float* dv, int64_t n,
                               int64_t* labels, uint64_t seed) {
  std::vector<int64_t> l(labels, labels + n);
  lp_outliers(csr_in(ip, ix, dv, n, n), l, seed);
  std::copy(l.begin(), l.end(), labels);
  return 0;
}

extern "C" int64_t orc_lp_remap(int64_t* labels, int64_t n) {
  std::vector<int64_t> l(labels, labels + n);
  const int64_t m = lp_remap(l);
  std::copy(l.begin(), l.end(), labels);
  return m;
}

extern "C" int64_t orc_lp_loop(const int64_t* ip, const int64_t* ix, const float* dv, int64_t n,
                               int64_t approx_parts, int n_iter, uint64_t seed, int64_t* part) {
  int64_t np = 0;
  std::vector<int64_t> p = lp_loop(csr_in(ip, ix, dv, n, n), approx_parts, n_iter, seed, &np);
  std::copy(p.begin(), p.end(), part);
  return np;
}

extern "C" int orc_unit_range(float* m, int64_t n, int64_t d) {
  std::vector<float> v(m, m + n * d);
  unit_range(v, n, d);
  std::copy(v.begin(), v.end(), m);
  return 0;
}

extern "C" int orc_pca_init(const float* data, int64_t n, int64_t d, int dim, float* out) {
  std::vector<float> v(data, data + n * d);
  std::vector<float> r = pca_init(v, n, d, dim);
  std::copy(r.begin(), r.end(), out);
  return 0;
}

extern "C" int orc_random_init(int64_t n, int dim, uint64_t seed, float* out) {
  std::vector<float> r = random_init(n, dim, seed);
  std::copy(r.begin(), r.end(), out);
  return 0;
}

// expand_layout on its own, the partition given as labels
extern "C" int orc_lp_expand(const int64_t* ip, const int64_t* ix, const float* dv, int64_t n,
                             const int64_t* part, int64_t n_parts, const float* layout, int dim,
                             double scaling, float* out) {
  Csr nrm = norm_cols_l2(from_partition(std::vector<int64_t>(part, part + n), n_parts));
  std::vector<float> lay(layout, layout + n_parts * dim);
  std::vector<float> r = expand_layout(csr_in(ip, ix, dv, n, n), nrm, lay, dim, scaling);
  std::copy(r.begin(), r.end(), out);
  return 0;
}

// ================================================================ A8/A9: core distances + MST
namespace {
// core_sq[i]: squared distance to the min(min_samples, n-1)-th nearest other point
// (src/clustering/mst.rs:92-105 with KdTree::knn_query_batch, kd_tree.rs:471-486).
template <typename T>
void core_sq_impl(const T* e, int64_t n, int dim, int min_samples, T* core) {
  const int64_t k = std::min<int64_t>(min_samples, n - 1);
  if (k <= 0) {
    for (int64_t i = 0; i < n; ++i) core[i] = T(0);
    return;
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    std::vector<T> best;
    best.reserve(k + 1);
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      T dd = dsq_seq(e + i * dim, e + j * dim, dim);
      if ((int64_t)best.size() == k && !(dd < best.back())) continue;
      best.insert(std::upper_bound(best.begin(), best.end(), dd), dd);
      if ((int64_t)best.size() > k) best.pop_back();
    }
    core[i] = best.back();
  }
}

// Prim's algorithm on the dense mutual-reachability graph under the strict total
// order (w^2, min(u,v), max(u,v)); this spanning tree is the unique minimum one for
// that order, hence the same edge set any exact Boruvka (src/clustering/mst.rs:82-166)
// produces when its ties are broken the same way.
template <typename T>
struct EKey {
  T w;
  int32_t a, b;
  bool operator<(const EKey& o) const {
    if (w != o.w) return w < o.w;
    if (a != o.a) return a < o.a;
    return b < o.b;
  }
};
template <typename T>
int64_t mst_impl(const T* e, int64_t n, int dim, int min_samples, int32_t* uo, int32_t* vo,
                 T* wo) {
  if (n <= 1) return 0;
  std::vector<T> core(n);
  core_sq_impl(e, n, dim, min_samples, core.data());
  std::vector<char> in_tree(n, 0);
  std::vector<EKey<T>> best(n);
  const T big = std::numeric_limits<T>::infinity();
  for (int64_t j = 0; j < n; ++j) best[j] = {big, INT32_MAX, INT32_MAX};
  std::vector<EKey<T>> edges;
  edges.reserve(n - 1);
  int64_t cur = 0;
  in_tree[0] = 1;
  for (int64_t step = 0; step < n - 1; ++step) {
    const T* ec = e + cur * dim;
    const T cc = core[cur];
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
      if (in_tree[j]) continue;
      T dd = dsq_seq(ec, e + j * dim, dim);
      T mr = std::max(std::max(dd, cc), core[j]);
      EKey<T> key{mr, (int32_t)std::min<int64_t>(cur, j), (int32_t)std::max<int64_t>(cur, j)};
      if (key < best[j]) best[j] = key;
    }
    int64_t arg = -1;
    for (int64_t j = 0; j < n; ++j)
      if (!in_tree[j] && (arg < 0 || best[j] < best[arg])) arg = j;
    edges.push_back(best[arg]);
    in_tree[arg] = 1;
    cur = arg;
  }
  std::sort(edges.begin(), edges.end());
  for (int64_t i = 0; i < n - 1; ++i) {
    uo[i] = edges[i].a;
    vo[i] = edges[i].b;
    wo[i] = std::sqrt(edges[i].w);
  }
  return n - 1;
}

// ---------------------------------------------------------------- reference complexity
// The same core distances and the same (unique) minimum spanning tree in O(n log n) work
// per round, in the shape of the reference's own algorithm: a kd-tree over the embedding
// (src/clustering/kd_tree.rs:93-214), k-nearest-neighbour queries for the core distances
// (kd_tree.rs:471-486) and Boruvka rounds of nearest-other-component queries with
// component pruning (kd_tree.rs:272-419, src/clustering/mst.rs:107-165). Used by the CPU
// baseline; the dense Prim above stays the parity reference, and the two must agree.
template <typename T>
struct KdTree {
  int dim = 0;
  int64_t n = 0;
  std::vector<int64_t> perm;      // position -> point
  std::vector<T> pts;             // coordinates by position
  struct Node { int64_t lo, hi; int32_t left, right; };
  std::vector<Node> nodes;
  std::vector<T> bmin, bmax;      // node boxes
  static constexpr int LEAF = 16;
  void build(const T* e, int64_t n_, int dim_) {
    n = n_;
    dim = dim_;
    perm.resize(n);
    std::iota(perm.begin(), perm.end(), 0);
    nodes.clear();
    bmin.clear();
    bmax.clear();
    std::vector<int32_t> stack{make(e, 0, n)};
    while (!stack.empty()) {
      const int32_t id = stack.back();
      stack.pop_back();
      const int64_t lo = nodes[id].lo, hi = nodes[id].hi;
      if (hi - lo <= LEAF) continue;
      int c = 0;
      T wmax = -1;
      for (int q = 0; q < dim; ++q) {
        T w = bmax[id * dim + q] - bmin[id * dim + q];
        if (w > wmax) { wmax = w; c = q; }
      }
      const int64_t mid = (lo + hi) / 2;
      std::nth_element(perm.begin() + lo, perm.begin() + mid, perm.begin() + hi,
                       [&](int64_t a, int64_t b) {
                         const T va = e[a * dim + c], vb = e[b * dim + c];
                         return va != vb ? va < vb : a < b;
                       });
      const int32_t l = make(e, lo, mid), r = make(e, mid, hi);
      nodes[id].left = l;
      nodes[id].right = r;
      stack.push_back(l);
      stack.push_back(r);
    }
    pts.resize(n * dim);
    for (int64_t q = 0; q < n; ++q)
      for (int c = 0; c < dim; ++c) pts[q * dim + c] = e[perm[q] * dim + c];
  }
  int32_t make(const T* e, int64_t lo, int64_t hi) {
    const int32_t id = (int32_t)nodes.size();
    nodes.push_back({lo, hi, -1, -1});
    bmin.resize((size_t)(id + 1) * dim, std::numeric_limits<T>::infinity());
    bmax.resize((size_t)(id + 1) * dim, -std::numeric_limits<T>::infinity());
    for (int64_t q = lo; q < hi; ++q)
      for (int c = 0; c < dim; ++c) {
        const T v = e[perm[q] * dim + c];
        bmin[id * dim + c] = std::min(bmin[id * dim + c], v);
        bmax[id * dim + c] = std::max(bmax[id * dim + c], v);
      }
    return id;
  }
  // a lower bound of dsq_seq(x, y) for every y in the box: each term's difference rounds
  // no larger than the point's own, and the sequential sum is monotone in every term
  T box_lb(const T* x, int32_t id) const {
    T s = T(0);
    for (int c = 0; c < dim; ++c) {
      T g = T(0);
      if (x[c] < bmin[id * dim + c]) g = bmin[id * dim + c] - x[c];
      else if (x[c] > bmax[id * dim + c]) g = x[c] - bmax[id * dim + c];
      T p = g * g;
      s = s + p;
    }
    return s;
  }
};

template <typename T>
void core_sq_kd(const T* e, int64_t n, int dim, int min_samples, T* core, const KdTree<T>& t,
                std::vector<int32_t>* nbrs) {
  const int64_t k = std::min<int64_t>(min_samples, n - 1);
  if (nbrs) nbrs->assign(n * std::max<int64_t>(k, 1), -1);
  if (k <= 0) {
    for (int64_t i = 0; i < n; ++i) core[i] = T(0);
    return;
  }
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    const T* x = e + i * dim;
    std::vector<std::pair<T, int64_t>> best;  // ascending (dsq, point)
    best.reserve(k + 1);
    std::vector<int32_t> st{0};
    while (!st.empty()) {
      const int32_t id = st.back();
      st.pop_back();
      if ((int64_t)best.size() == k && t.box_lb(x, id) > best.back().first) continue;
      const auto& nd = t.nodes[id];
      if (nd.left < 0) {
        for (int64_t q = nd.lo; q < nd.hi; ++q) {
          const int64_t j = t.perm[q];
          if (j == i) continue;
          const T dd = dsq_seq(x, &t.pts[q * dim], dim);
          std::pair<T, int64_t> c(dd, j);
          if ((int64_t)best.size() == k && !(c < best.back())) continue;
          best.insert(std::upper_bound(best.begin(), best.end(), c), c);
          if ((int64_t)best.size() > k) best.pop_back();
        }
        continue;
      }
      const T ll = t.box_lb(x, nd.left), lr = t.box_lb(x, nd.right);
      if (ll <= lr) { st.push_back(nd.right); st.push_back(nd.left); }
      else { st.push_back(nd.left); st.push_back(nd.right); }
    }
    core[i] = best.back().first;
    if (nbrs)
      for (int64_t q = 0; q < k; ++q) (*nbrs)[i * k + q] = (int32_t)best[q].second;
  }
}

template <typename T>
int64_t mst_boruvka_impl(const T* e, int64_t n, int dim, int min_samples, int32_t* uo,
                         int32_t* vo, T* wo) {
  if (n <= 1) return 0;
  KdTree<T> t;
  t.build(e, n, dim);
  std::vector<T> core(n);
  std::vector<int32_t> nb;
  core_sq_kd(e, n, dim, min_samples, core.data(), t, &nb);
  const int64_t kn = std::max<int64_t>(std::min<int64_t>(min_samples, n - 1), 1);
  const int32_t nn = (int32_t)t.nodes.size();
  // smallest core distance in each node (children after parents in `nodes`)
  std::vector<T> ncore(nn);
  for (int32_t id = nn - 1; id >= 0; --id) {
    const auto& nd = t.nodes[id];
    if (nd.left < 0) {
      T m = std::numeric_limits<T>::infinity();
      for (int64_t q = nd.lo; q < nd.hi; ++q) m = std::min(m, core[t.perm[q]]);
      ncore[id] = m;
    } else {
      ncore[id] = std::min(ncore[nd.left], ncore[nd.right]);
    }
  }
  std::vector<int64_t> parent(n);
  std::iota(parent.begin(), parent.end(), 0);
  auto find = [&](int64_t a) {
    while (parent[a] != a) { parent[a] = parent[parent[a]]; a = parent[a]; }
    return a;
  };
  std::vector<int64_t> comp(n), ncomp(nn);
  std::vector<EKey<T>> edges;
  edges.reserve(n - 1);
  const T inf = std::numeric_limits<T>::infinity();
  const EKey<T> none{inf, INT32_MAX, INT32_MAX};
  while ((int64_t)edges.size() < n - 1) {
    for (int64_t i = 0; i < n; ++i) comp[i] = find(i);
    for (int32_t id = nn - 1; id >= 0; --id) {
      const auto& nd = t.nodes[id];
      if (nd.left < 0) {
        int64_t c = comp[t.perm[nd.lo]];
        for (int64_t q = nd.lo + 1; q < nd.hi && c >= 0; ++q)
          if (comp[t.perm[q]] != c) c = -1;
        ncomp[id] = c;
      } else {
        ncomp[id] = ncomp[nd.left] == ncomp[nd.right] ? ncomp[nd.left] : -1;
      }
    }
    std::vector<EKey<T>> best(n, none);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
      const int64_t ci = comp[i];
      const T* x = e + i * dim;
      EKey<T> b = none;
      // the core-distance neighbours give a first bound
      for (int64_t q = 0; q < kn; ++q) {
        const int64_t j = nb[i * kn + q];
        if (j < 0 || comp[j] == ci) continue;
        const T mr = std::max(std::max(dsq_seq(x, e + j * dim, dim), core[i]), core[j]);
        EKey<T> key{mr, (int32_t)std::min(i, j), (int32_t)std::max(i, j)};
        if (key < b) b = key;
      }
      std::vector<int32_t> st{0};
      while (!st.empty()) {
        const int32_t id = st.back();
        st.pop_back();
        if (ncomp[id] == ci) continue;
        const T lb = std::max(std::max(t.box_lb(x, id), core[i]), ncore[id]);
        if (lb > b.w) continue;
        const auto& nd = t.nodes[id];
        if (nd.left < 0) {
          for (int64_t q = nd.lo; q < nd.hi; ++q) {
            const int64_t j = t.perm[q];
            if (comp[j] == ci) continue;
            const T mr = std::max(std::max(dsq_seq(x, &t.pts[q * dim], dim), core[i]), core[j]);
            EKey<T> key{mr, (int32_t)std::min(i, j), (int32_t)std::max(i, j)};
            if (key < b) b = key;
          }
          continue;
        }
        const T ll = t.box_lb(x, nd.left), lr = t.box_lb(x, nd.right);
        if (ll <= lr) { st.push_back(nd.right); st.push_back(nd.left); }
        else { st.push_back(nd.left); st.push_back(nd.right); }
      }
      best[i] = b;
    }
    std::vector<EKey<T>> cbest(n, none);
    for (int64_t i = 0; i < n; ++i)
      if (best[i] < cbest[comp[i]]) cbest[comp[i]] = best[i];
    int64_t added = 0;
    for (int64_t c = 0; c < n; ++c) {
      if (comp[c] != c || cbest[c].a == INT32_MAX) continue;
      const int64_t ra = find(cbest[c].a), rb = find(cbest[c].b);
      if (ra == rb) continue;  // the other component chose the same edge
      parent[ra] = rb;
      edges.push_back(cbest[c]);
      ++added;
    }
    if (added == 0) break;  // disconnected (not for finite input)
  }
  std::sort(edges.begin(), edges.end());
  const int64_t ne = (int64_t)edges.size();
  for (int64_t i = 0; i < ne; ++i) {
    uo[i] = edges[i].a;
    vo[i] = edges[i].b;
    wo[i] = std::sqrt(edges[i].w);
  }
  return ne;
}

// which MST the layer extraction uses: 0 the dense Prim (parity reference), 1 the kd-tree
// Boruvka (same tree, reference complexity)
int g_mst_algo = 0;
template <typename T>
int64_t mst_any(const T* e, int64_t n, int dim, int ms, int32_t* u, int32_t* v, T* w) {
  return g_mst_algo == 1 ? mst_boruvka_impl<T>(e, n, dim, ms, u, v, w)
                         : mst_impl<T>(e, n, dim, ms, u, v, w);
}

// mst_to_linkage_tree: src/clustering/linkage.rs:323-367. The stable sort by weight is
// made total with (w, min(u,v), max(u,v)); the parent walk is a union-find whose
// roots carry the current cluster label (same result, no O(n^2) walk).
template <typename T>
int linkage_impl(int32_t* u, int32_t* v, T* w, int64_t ne, int64_t n, int64_t* left,
                 int64_t* right, T* dist, int64_t* size) {
  std::vector<int64_t> order(ne);
  std::iota(order.begin(), order.end(), 0);
  auto key = [&](int64_t i) {
    return std::make_tuple(w[i], std::min(u[i], v[i]), std::max(u[i], v[i]));
  };
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return key(a) < key(b); });
  std::vector<int32_t> su(ne), sv(ne);
  std::vector<T> sw(ne);
  for (int64_t i = 0; i < ne; ++i) {
    su[i] = u[order[i]];
    sv[i] = v[order[i]];
    sw[i] = w[order[i]];
  }
  if (ne > 0) {  // an empty vector's data() may be null, which memcpy does not accept
    std::memcpy(u, su.data(), ne * sizeof(int32_t));
    std::memcpy(v, sv.data(), ne * sizeof(int32_t));
    std::memcpy(w, sw.data(), ne * sizeof(T));
  }
  std::vector<int64_t> par(n), lab(n), sz(2 * n, 0);
  std::iota(par.begin(), par.end(), 0);
  std::iota(lab.begin(), lab.end(), 0);
  for (int64_t i = 0; i < n; ++i) sz[i] = 1;
  auto find = [&](int64_t x) {
    while (par[x] != x) {
      par[x] = par[par[x]];
      x = par[x];
    }
    return x;
  };
  for (int64_t i = 0; i < ne; ++i) {
    int64_t ru = find(u[i]), rv = find(v[i]);
    int64_t l = lab[ru], r = lab[rv];
    int64_t ns = sz[l] + sz[r];
    if (l < r) std::swap(l, r);
    left[i] = l;
    right[i] = r;
    dist[i] = w[i];
    size[i] = ns;
    sz[n + i] = ns;
    par[ru] = rv;
    lab[rv] = n + i;
  }
  return 0;
}

// condense_tree: src/clustering/condensed_tree.rs:127-248 (with collect_descendants
// :54-70 and eliminate_branch :86-107)
template <typename T>
struct CNode {
  int64_t parent, child;
  T lambda;
  int64_t size;
};
template <typename T>
std::vector<CNode<T>> condense_impl(const int64_t* left, const int64_t* right, const T* dist,
                                    const int64_t* size, int64_t nrows, int64_t n, int64_t mcs) {
  std::vector<CNode<T>> out;
  if (nrows == 0) return out;
  const int64_t root = 2 * nrows;
  auto descend = [&](int64_t r, std::vector<int64_t>& res) {
    res.clear();
    std::vector<int64_t> st{r};
    while (!st.empty()) {
      int64_t x = st.back();
      st.pop_back();
      res.push_back(x);
      if (x >= n) {
        st.push_back(left[x - n]);
        st.push_back(right[x - n]);
      }
    }
  };
  std::vector<int64_t> trav;
  descend(root, trav);
  std::vector<int64_t> relabel(root + 1, 0);
  relabel[root] = n;
  int64_t next = n + 1;
  std::vector<char> ignore(root + 1, 0);
  std::vector<int64_t> tmp;
  auto eliminate = [&](int64_t node, int64_t pl, T lam) {
    descend(node, tmp);
    for (int64_t x : tmp) {
      if (x < n) out.push_back({pl, x, lam, 1});
      else ignore[x] = 1;
    }
  };
  for (int64_t node : trav) {
    if (node < n || ignore[node]) continue;
    int64_t pl = relabel[node];
    int64_t l = left[node - n], r = right[node - n];
    T d = dist[node - n];
    T lam = d > T(0) ? T(1) / d : std::numeric_limits<T>::infinity();
    int64_t ls = l >= n ? size[l - n] : 1, rs = r >= n ? size[r - n] : 1;
    bool lb = ls >= mcs, rb = rs >= mcs;
    if (lb && rb) {
      relabel[l] = next;
      out.push_back({pl, next, lam, ls});
      ++next;
      relabel[r] = next;
      out.push_back({pl, next, lam, rs});
      ++next;
    } else if (lb && !rb) {
      relabel[l] = pl;
      eliminate(r, pl, lam);
    } else if (!lb && rb) {
      relabel[r] = pl;
      eliminate(l, pl, lam);
    } else {
      eliminate(l, pl, lam);
      eliminate(r, pl, lam);
    }
  }
  return out;
}

// extract_leaves: src/clustering/condensed_tree.rs:264-294
template <typename T>
std::vector<int64_t> leaves_impl(const std::vector<CNode<T>>& t, int64_t n) {
  std::vector<int64_t> res;
  if (t.empty()) return res;
  int64_t max_id = 0;
  for (auto& c : t) max_id = std::max(max_id, std::max(c.parent, c.child));
  std::vector<char> is_parent(max_id + 1, 0), appears(max_id + 1, 0);
  for (auto& c : t) {
    if (c.size > 1) is_parent[c.parent] = 1;
    appears[c.parent] = 1;
  }
  for (int64_t i = n; i <= max_id; ++i)
    if (!is_parent[i] && appears[i]) res.push_back(i);
  return res;
}

// DisjointSet (src/utils/disjoint_set.rs:12-98): union by rank, path halving
struct Dsu {
  std::vector<int64_t> p, rk;
  explicit Dsu(int64_t n) : p(n), rk(n, 0) { std::iota(p.begin(), p.end(), 0); }
  int64_t find(int64_t x) {
    while (p[x] != x) {
      p[x] = p[p[x]];
      x = p[x];
    }
    return x;
  }
  bool unite(int64_t x, int64_t y) {
    int64_t rx = find(x), ry = find(y);
    if (rx == ry) return false;
    if (rk[rx] < rk[ry]) p[rx] = ry;
    else if (rk[rx] > rk[ry]) p[ry] = rx;
    else {
      p[ry] = rx;
      rk[rx]++;
    }
    return true;
  }
};

// get_cluster_label_vector: src/clustering/condensed_tree.rs:311-369
template <typename T>
std::vector<int64_t> labels_impl(const std::vector<CNode<T>>& t,
                                 const std::vector<int64_t>& clusters, int64_t n) {
  std::vector<int64_t> labels(n, -1);
  if (clusters.empty()) return labels;
  int64_t max_id = 0, root = std::numeric_limits<int64_t>::max();
  for (auto& c : t) {
    max_id = std::max(max_id, std::max(c.parent, c.child));
    root = std::min(root, c.parent);
  }
  if (t.empty()) root = n;
  Dsu ds(max_id + 1);
  std::unordered_set<int64_t> cs(clusters.begin(), clusters.end());
  for (auto& c : t)
    if (!cs.count(c.child)) ds.unite(c.parent, c.child);
  std::vector<int64_t> sorted(clusters);
  std::sort(sorted.begin(), sorted.end());
  std::unordered_map<int64_t, int64_t> lm;
  for (size_t i = 0; i < sorted.size(); ++i) lm[sorted[i]] = (int64_t)i;
  for (int64_t i = 0; i < n; ++i) {
    if (i > max_id) continue;
    int64_t c = ds.find(i);
    if (c <= root) labels[i] = -1;
    else {
      auto it = lm.find(c);
      labels[i] = it == lm.end() ? -1 : it->second;
    }
  }
  return labels;
}

// get_point_membership_strengths + max_lambdas: condensed_tree.rs:384-484
template <typename T>
std::vector<T> strengths_impl(const std::vector<CNode<T>>& t, const std::vector<int64_t>& clusters,
                              const std::vector<int64_t>& labels) {
  const int64_t n = (int64_t)labels.size();
  std::vector<T> res(n, T(0));
  std::unordered_map<int64_t, T> deaths;
  for (auto c : clusters) deaths[c] = T(0);
  for (auto& c : t)
    if (c.size == 1) {
      auto it = deaths.find(c.parent);
      if (it != deaths.end() && c.lambda > it->second) it->second = c.lambda;
    }
  std::vector<int64_t> sorted(clusters);
  std::sort(sorted.begin(), sorted.end());
  int64_t root = n;
  if (!t.empty()) {
    root = std::numeric_limits<int64_t>::max();
    for (auto& c : t) root = std::min(root, c.parent);
  }
  for (auto& c : t) {
    int64_t pt = c.child;
    if (pt >= root || pt >= n) continue;
    if (labels[pt] < 0) continue;
    if (labels[pt] >= (int64_t)sorted.size()) continue;
    int64_t cid = sorted[labels[pt]];
    auto it = deaths.find(cid);
    if (it == deaths.end()) continue;
    T ml = it->second;
    if (ml == T(0) || !std::isfinite(c.lambda)) res[pt] = T(1);
    else {
      T lv = c.lambda < ml ? c.lambda : ml;
      res[pt] = lv / ml;
    }
  }
  return res;
}

// min_cluster_size_barcode + compute_total_persistence: persistence.rs:42-173, with the
// barcode arithmetic in P (float = reference precision, double = fp64 variant)
template <typename P, typename T>
struct Barcode {
  P birth, death;
  int64_t parent;
  P lambda_death;
};
template <typename P, typename T>
std::vector<Barcode<P, T>> barcode_impl(const std::vector<CNode<T>>& t, int64_t n,
                                        int64_t min_size) {
  std::vector<Barcode<P, T>> bc;
  if (t.empty()) return bc;
  std::vector<const CNode<T>*> cn;
  for (auto& c : t)
    if (c.size > 1) cn.push_back(&c);
  if (cn.empty()) return bc;
  int64_t max_child = 0;
  for (auto* c : cn) max_child = std::max(max_child, c->child);
  int64_t nn = max_child - n + 1;
  bc.assign(nn, {(P)min_size, P(0), n, P(0)});
  bc[0].lambda_death = P(0);
  bc[0].death = (P)n;
  int64_t i = (int64_t)cn.size();
  while (i >= 2) {
    i -= 2;
    const CNode<T>* a = cn[i];
    const CNode<T>* b = cn[i + 1];
    int64_t ia = a->child - n, ib = b->child - n;
    bc[ia].parent = a->parent;
    bc[ib].parent = b->parent;
    P lv = std::exp(-P(1) / (P)a->lambda);
    bc[ia].lambda_death = lv;
    bc[ib].lambda_death = lv;
    P death_t = (P)std::min(a->size, b->size);
    bc[ia].death = death_t;
    bc[ib].death = death_t;
    int64_t pidx = a->parent - n;
    P mb = std::max(std::max(bc[ia].birth, bc[ib].birth), death_t);
    bc[pidx].birth = std::max(bc[pidx].birth, mb);
  }
  return bc;
}
template <typename P, typename T>
void total_persistence(const std::vector<Barcode<P, T>>& bc, std::vector<P>& sizes,
                       std::vector<P>& pers) {
  sizes.clear();
  pers.clear();
  if (bc.empty()) return;
  for (auto& b : bc) sizes.push_back(b.birth);
  std::sort(sizes.begin(), sizes.end());
  sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
  pers.assign(sizes.size(), P(0));
  for (size_t bi = 1; bi < bc.size(); ++bi) {
    const auto& b = bc[bi];
    if (b.death <= b.birth) continue;
    size_t bp = std::lower_bound(sizes.begin(), sizes.end(), b.birth) - sizes.begin();
    size_t dp = std::lower_bound(sizes.begin(), sizes.end(), b.death) - sizes.begin();
    P contrib = (b.death - b.birth) * b.lambda_death;
    for (size_t k = bp; k < dp; ++k) pers[k] += contrib;
  }
}
// find_peaks: persistence.rs:188-214
template <typename P>
std::vector<int64_t> peaks_impl(const std::vector<P>& x) {
  std::vector<int64_t> pk;
  if (x.size() < 3) return pk;
  size_t i = 1, imax = x.size() - 1;
  while (i < imax) {
    if (x[i - 1] < x[i]) {
      size_t ia = i + 1;
      while (ia < imax && x[ia] == x[i]) ++ia;
      if (x[ia] < x[i]) {
        pk.push_back((int64_t)((i + ia - 1) / 2));
        i = ia;
      }
    }
    ++i;
  }
  return pk;
}
// estimate_cluster_similarity + select_diverse_peaks: persistence.rs:231-318
template <typename P>
double similarity(const std::vector<P>& birth, const std::vector<P>& death, P a, P b) {
  int64_t inter = 0, uni = 0;
  for (size_t i = 0; i < birth.size(); ++i) {
    bool ia = birth[i] <= a && death[i] > a;
    bool ib = birth[i] <= b && death[i] > b;
    inter += (ia && ib);
    uni += (ia || ib);
  }
  return uni == 0 ? 0.0 : (double)inter / (double)uni;
}
template <typename P>
std::vector<int64_t> select_peaks(const std::vector<int64_t>& peaks, const std::vector<P>& pers,
                                  const std::vector<P>& sizes, const std::vector<P>& birth,
                                  const std::vector<P>& death, double min_sim,
                                  int64_t max_layers) {
  std::vector<int64_t> sel;
  if (peaks.empty()) return sel;
  std::vector<std::pair<int64_t, P>> idx;
  for (auto p : peaks) idx.push_back({p, pers[p]});
  std::stable_sort(idx.begin(), idx.end(),
                   [](const std::pair<int64_t, P>& a, const std::pair<int64_t, P>& b) {
                     return b.second < a.second;
                   });
  std::vector<P> sb;
  for (auto& pp : idx) {
    if ((int64_t)sel.size() >= max_layers) break;
    P bs = sizes[pp.first];
    bool diverse = true;
    for (P s : sb)
      if (!(similarity(birth, death, bs, s) <= min_sim)) {
        diverse = false;
        break;
      }
    if (diverse) {
      sel.push_back(pp.first);
      sb.push_back(bs);
    }
  }
  return sel;
}

int64_t n_clusters_of(const std::vector<int64_t>& l) {
  int64_t m = -1;
  for (auto x : l)
    if (x >= 0) m = std::max(m, x);
  return m + 1;
}

// build_cluster_layers: persistence.rs:387-463 (P = persistence arithmetic type)
template <typename P, typename T>
int64_t layers_impl(const T* emb, int64_t n, int dim, int min_samples, int64_t base_mcs,
                    double min_sim, int64_t max_layers, int64_t* labels_out, T* str_out,
                    double* pers_out) {
  if (n == 0) return 0;
  std::vector<int32_t> u(n), v(n);
  std::vector<T> w(n);
  int64_t ne = mst_any<T>(emb, n, dim, min_samples, u.data(), v.data(), w.data());
  std::vector<int64_t> L(ne), R(ne), S(ne);
  std::vector<T> D(ne);
  linkage_impl<T>(u.data(), v.data(), w.data(), ne, n, L.data(), R.data(), D.data(), S.data());
  auto ct = condense_impl<T>(L.data(), R.data(), D.data(), S.data(), ne, n, base_mcs);
  auto leaves = leaves_impl(ct, n);
  std::vector<std::vector<int64_t>> all_l{labels_impl(ct, leaves, n)};
  std::vector<std::vector<T>> all_s{strengths_impl(ct, leaves, all_l[0])};
  std::vector<double> all_p{0.0};
  auto bc = barcode_impl<P, T>(ct, n, base_mcs);
  if (!bc.empty()) {
    std::vector<P> sizes, pers;
    total_persistence(bc, sizes, pers);
    auto pk = peaks_impl(pers);
    std::vector<P> birth, death;
    for (auto& b : bc) {
      birth.push_back(b.birth);
      death.push_back(b.death);
    }
    int64_t ml = max_layers > 0 ? max_layers - 1 : 0;
    auto sel = select_peaks(pk, pers, sizes, birth, death, min_sim, ml);
    for (auto p : sel) {
      P bs = sizes[p];
      std::vector<int64_t> act;
      for (size_t i = 0; i < bc.size(); ++i)
        if (bc[i].birth <= bs && bc[i].death > bs) act.push_back((int64_t)i + n);
      all_l.push_back(labels_impl(ct, act, n));
      all_s.push_back(strengths_impl(ct, act, all_l.back()));
      all_p.push_back((double)pers[p]);
    }
  }
  std::vector<int64_t> order(all_l.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    return n_clusters_of(all_l[b]) < n_clusters_of(all_l[a]);
  });
  int64_t nl = std::min<int64_t>((int64_t)order.size(), std::max<int64_t>(max_layers, 1));
  for (int64_t li = 0; li < nl; ++li) {
    int64_t o = order[li];
    std::memcpy(labels_out + li * n, all_l[o].data(), n * sizeof(int64_t));
    std::memcpy(str_out + li * n, all_s[o].data(), n * sizeof(T));
    pers_out[li] = all_p[o];
  }
  return nl;
}

// search_for_n_clusters: src/lib.rs:377-447
template <typename T>
int search_impl(const T* emb, int64_t n, int dim, int min_samples, int64_t target_k,
                int64_t* labels_out, T* str_out) {
  if (n == 0) return 0;
  std::vector<int32_t> u(n), v(n);
  std::vector<T> w(n);
  int64_t ne = mst_any<T>(emb, n, dim, min_samples, u.data(), v.data(), w.data());
  std::vector<int64_t> L(ne), R(ne), S(ne);
  std::vector<T> D(ne);
  linkage_impl<T>(u.data(), v.data(), w.data(), ne, n, L.data(), R.data(), D.data(), S.data());
  int64_t lo = 2, hi = n / 2;
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) / 2;
    if (mid == lo || mid == hi) break;
    auto ct = condense_impl<T>(L.data(), R.data(), D.data(), S.data(), ne, n, mid);
    int64_t mk = (int64_t)leaves_impl(ct, n).size();
    if (mk < target_k) hi = mid;
    else lo = mid;
  }
  auto ctl = condense_impl<T>(L.data(), R.data(), D.data(), S.data(), ne, n, lo);
  auto lvl = leaves_impl(ctl, n);
  auto labl = labels_impl(ctl, lvl, n);
  auto cth = condense_impl<T>(L.data(), R.data(), D.data(), S.data(), ne, n, hi);
  auto lvh = leaves_impl(cth, n);
  auto labh = labels_impl(cth, lvh, n);
  int64_t ld = std::llabs((int64_t)lvl.size() - target_k);
  int64_t hd = std::llabs((int64_t)lvh.size() - target_k);
  bool use_lo;
  if (ld < hd) use_lo = true;
  else if (hd < ld) use_lo = false;
  else {
    int64_t la = 0, ha = 0;
    for (auto x : labl) la += x >= 0;
    for (auto x : labh) ha += x >= 0;
    use_lo = la >= ha;
  }
  std::vector<T> s = use_lo ? strengths_impl(ctl, lvl, labl) : strengths_impl(cth, lvh, labh);
  std::memcpy(labels_out, (use_lo ? labl : labh).data(), n * sizeof(int64_t));
  std::memcpy(str_out, s.data(), n * sizeof(T));
  return 0;
}

template <typename T>
std::vector<CNode<T>> nodes_from(const int64_t* parent, const int64_t* child, const T* lam,
                                 const int64_t* cs, int64_t nn) {
  std::vector<CNode<T>> t(nn);
  for (int64_t i = 0; i < nn; ++i) t[i] = {parent[i], child[i], lam ? lam[i] : T(0), cs ? cs[i] : 1};
  return t;
}
}  // namespace

extern "C" int orc_core_sq(const float* emb, int64_t n, int dim, int ms, float* core) {
  core_sq_impl<float>(emb, n, dim, ms, core);
  return 0;
}
extern "C" int orc_core_sq_f64(const double* emb, int64_t n, int dim, int ms, double* core) {
  core_sq_impl<double>(emb, n, dim, ms, core);
  return 0;
}
extern "C" int64_t orc_mst(const float* emb, int64_t n, int dim, int ms, int32_t* u, int32_t* v,
                           float* w) {
  return mst_impl<float>(emb, n, dim, ms, u, v, w);
}
extern "C" int64_t orc_mst_f64(const double* emb, int64_t n, int dim, int ms, int32_t* u,
                               int32_t* v, double* w) {
  return mst_impl<double>(emb, n, dim, ms, u, v, w);
}
extern "C" int64_t orc_mst_boruvka(const float* emb, int64_t n, int dim, int ms, int32_t* u,
                                   int32_t* v, float* w) {
  return mst_boruvka_impl<float>(emb, n, dim, ms, u, v, w);
}
extern "C" int orc_core_sq_kd(const float* emb, int64_t n, int dim, int ms, float* core) {
  if (n == 0) return 0;
  KdTree<float> t;
  t.build(emb, n, dim);
  core_sq_kd<float>(emb, n, dim, ms, core, t, nullptr);
  return 0;
}
extern "C" int orc_set_mst_algo(int algo) {
  const int old = g_mst_algo;
  g_mst_algo = algo;
  return old;
}

extern "C" int orc_linkage(int32_t* u, int32_t* v, float* w, int64_t ne, int64_t n,
                           int64_t* left, int64_t* right, float* dist, int64_t* size) {
  return linkage_impl<float>(u, v, w, ne, n, left, right, dist, size);
}
extern "C" int orc_linkage_f64(int32_t* u, int32_t* v, double* w, int64_t ne, int64_t n,
                               int64_t* left, int64_t* right, double* dist, int64_t* size) {
  return linkage_impl<double>(u, v, w, ne, n, left, right, dist, size);
}
template <typename T>
static int64_t condense_export(const int64_t* l, const int64_t* r, const T* d, const int64_t* s,
                               int64_t nrows, int64_t n, int64_t mcs, int64_t* parent,
                               int64_t* child, T* lam, int64_t* cs) {
  auto t = condense_impl<T>(l, r, d, s, nrows, n, mcs);
  for (size_t i = 0; i < t.size(); ++i) {
    parent[i] = t[i].parent;
    child[i] = t[i].child;
    lam[i] = t[i].lambda;
    cs[i] = t[i].size;
  }
  return (int64_t)t.size();
}
extern "C" int64_t orc_condense(const int64_t* l, const int64_t* r, const float* d,
                                const int64_t* s, int64_t nrows, int64_t n, int64_t mcs,
                                int64_t* parent, int64_t* child, float* lam, int64_t* cs) {
  return condense_export<float>(l, r, d, s, nrows, n, mcs, parent, child, lam, cs);
}
extern "C" int64_t orc_condense_f64(const int64_t* l, const int64_t* r, const double* d,
                                    const int64_t* s, int64_t nrows, int64_t n, int64_t mcs,
                                    int64_t* parent, int64_t* child, double* lam, int64_t* cs) {
  return condense_export<double>(l, r, d, s, nrows, n, mcs, parent, child, lam, cs);
}
extern "C" int64_t orc_extract_leaves(const int64_t* parent, const int64_t* child,
                                      const int64_t* cs, int64_t nn, int64_t n, int64_t* out) {
  auto t = nodes_from<float>(parent, child, nullptr, cs, nn);
  auto lv = leaves_impl(t, n);
  for (size_t i = 0; i < lv.size(); ++i) out[i] = lv[i];
  return (int64_t)lv.size();
}
extern "C" int orc_label_vector(const int64_t* parent, const int64_t* child, int64_t nn,
                                const int64_t* clusters, int64_t nc, int64_t n, int64_t* labels) {
  auto t = nodes_from<float>(parent, child, nullptr, nullptr, nn);
  auto l = labels_impl(t, std::vector<int64_t>(clusters, clusters + nc), n);
  std::memcpy(labels, l.data(), n * sizeof(int64_t));
  return 0;
}
extern "C" int orc_strengths(const int64_t* parent, const int64_t* child, const float* lam,
                             const int64_t* cs, int64_t nn, const int64_t* clusters, int64_t nc,
                             const int64_t* labels, int64_t n, float* out) {
  auto t = nodes_from<float>(parent, child, lam, cs, nn);
  auto s = strengths_impl(t, std::vector<int64_t>(clusters, clusters + nc),
                          std::vector<int64_t>(labels, labels + n));
  std::memcpy(out, s.data(), n * sizeof(float));
  return 0;
}
extern "C" int64_t orc_find_peaks(const double* x, int64_t len, int64_t* peaks) {
  auto p = peaks_impl(std::vector<double>(x, x + len));
  for (size_t i = 0; i < p.size(); ++i) peaks[i] = p[i];
  return (int64_t)p.size();
}
template <typename P>
static int64_t persistence_export(const std::vector<CNode<float>>& t, int64_t n, int64_t mcs,
                                  double* sizes_out, double* pers_out, double* b_out,
                                  double* d_out, double* l_out) {
  auto bc = barcode_impl<P, float>(t, n, mcs);
  std::vector<P> sizes, pers;
  total_persistence(bc, sizes, pers);
  for (size_t i = 0; i < sizes.size(); ++i) {
    sizes_out[i] = (double)sizes[i];
    pers_out[i] = (double)pers[i];
  }
  for (size_t i = 0; i < bc.size(); ++i) {
    if (b_out) b_out[i] = (double)bc[i].birth;
    if (d_out) d_out[i] = (double)bc[i].death;
    if (l_out) l_out[i] = (double)bc[i].lambda_death;
  }
  return ((int64_t)bc.size() << 32) | (int64_t)sizes.size();
}
extern "C" int64_t orc_persistence(const int64_t* parent, const int64_t* child, const float* lam,
                                   const int64_t* cs, int64_t nn, int64_t n, int64_t mcs,
                                   int fp64, double* sizes_out, double* pers_out, double* b_out,
                                   double* d_out, double* l_out) {
  auto t = nodes_from<float>(parent, child, lam, cs, nn);
  if (fp64) return persistence_export<double>(t, n, mcs, sizes_out, pers_out, b_out, d_out, l_out);
  return persistence_export<float>(t, n, mcs, sizes_out, pers_out, b_out, d_out, l_out);
}
extern "C" int orc_select_diverse_peaks(const int64_t* peaks, int64_t np, const double* pers,
                                        const double* sizes, const double* births,
                                        const double* deaths, int64_t nbc, double min_sim,
                                        int64_t max_layers, int64_t* selected, int64_t* nsel) {
  int64_t ns = 0;
  for (int64_t i = 0; i < np; ++i) ns = std::max(ns, peaks[i] + 1);
  std::vector<double> P(pers, pers + ns), S(sizes, sizes + ns), B(births, births + nbc),
      D(deaths, deaths + nbc);
  auto s = select_peaks(std::vector<int64_t>(peaks, peaks + np), P, S, B, D, min_sim, max_layers);
  for (size_t i = 0; i < s.size(); ++i) selected[i] = s[i];
  *nsel = (int64_t)s.size();
  return 0;
}
extern "C" int64_t orc_cluster_layers(const float* emb, int64_t n, int dim, int ms, int64_t mcs,
                                      double min_sim, int64_t max_layers, int fp64,
                                      int64_t* labels, float* strengths, double* pers) {
  if (fp64)
    return layers_impl<double, float>(emb, n, dim, ms, mcs, min_sim, max_layers, labels,
                                      strengths, pers);
  return layers_impl<float, float>(emb, n, dim, ms, mcs, min_sim, max_layers, labels, strengths,
                                   pers);
}
extern "C" int64_t orc_cluster_layers_f64(const double* emb, int64_t n, int dim, int ms,
                                          int64_t mcs, double min_sim, int64_t max_layers,
                                          int64_t* labels, double* strengths, double* pers) {
  return layers_impl<double, double>(emb, n, dim, ms, mcs, min_sim, max_layers, labels,
                                     strengths, pers);
}
extern "C" int orc_search_n_clusters(const float* emb, int64_t n, int dim, int ms,
                                     int64_t target, int64_t* labels, float* strengths) {
  return search_impl<float>(emb, n, dim, ms, target, labels, strengths);
}
extern "C" int orc_search_n_clusters_f64(const double* emb, int64_t n, int dim, int ms,
                                         int64_t target, int64_t* labels, double* strengths) {
  return search_impl<double>(emb, n, dim, ms, target, labels, strengths);
}

// ================================================================ A16: evoc()
// src/lib.rs:199-352
extern "C" int64_t orc_evoc(const float* x, int64_t n, int64_t d, const int32_t* pre_idx,
                            const float* pre_dist, const orc_params_t* p, uint64_t seed,
                            int64_t* labels, float* strengths, double* persistence,
                            int32_t* knn_idx, float* knn_dist, float* embedding, float* init_out,
                            double* timings) {
  using clk = std::chrono::steady_clock;
  auto secs = [](clk::time_point a) {
    return std::chrono::duration<double>(clk::now() - a).count();
  };
  const int k = p->n_neighbours;
  std::vector<int32_t> idx(n * k);
  std::vector<float> dist(n * k);
  auto t0 = clk::now();
  if (pre_idx && pre_dist) {
    std::memcpy(idx.data(), pre_idx, n * k * sizeof(int32_t));
    std::memcpy(dist.data(), pre_dist, n * k * sizeof(float));
  } else {
    orc_knn_exhaustive(x, n, d, p->metric, k, idx.data(), dist.data());
  }
  if (timings) timings[0] = secs(t0);
  if (knn_idx) std::memcpy(knn_idx, idx.data(), n * k * sizeof(int32_t));
  if (knn_dist) std::memcpy(knn_dist, dist.data(), n * k * sizeof(float));

  t0 = clk::now();
  float eff_k = p->neighbour_scale * (float)k;
  std::vector<int64_t> indptr(n + 1);
  std::vector<int32_t> ind(2 * n * k);
  std::vector<float> wts(2 * n * k);
  orc_fuzzy_graph(idx.data(), dist.data(), n, k, eff_k, p->symmetrise, nullptr, nullptr,
                  indptr.data(), ind.data(), wts.data());
  if (timings) timings[1] = secs(t0);
  int dim = p->embedding_dim > 0 ? p->embedding_dim : std::min(std::max(k / 4, 4), 16);

  t0 = clk::now();
  std::vector<float> init(n * dim);
  // Csr::from_coo sorts each row by column (src/utils/sparse.rs:195-235)
  if (p->symmetrise) {
    orc_label_prop_init(indptr.data(), ind.data(), wts.data(), n, dim, x, d, seed, init.data());
  } else {
    std::vector<int64_t> ip2(indptr);
    std::vector<int32_t> i2(ind.begin(), ind.begin() + indptr[n]);
    std::vector<float> w2(wts.begin(), wts.begin() + indptr[n]);
    for (int64_t i = 0; i < n; ++i) {
      std::vector<std::pair<int32_t, float>> row;
      for (int64_t q = indptr[i]; q < indptr[i + 1]; ++q) row.push_back({ind[q], wts[q]});
      std::sort(row.begin(), row.end(),
                [](const std::pair<int32_t, float>& a, const std::pair<int32_t, float>& b) {
                  return a.first < b.first;
                });
      for (size_t q = 0; q < row.size(); ++q) {
        i2[indptr[i] + q] = row[q].first;
        w2[indptr[i] + q] = row[q].second;
      }
    }
    orc_label_prop_init(ip2.data(), i2.data(), w2.data(), n, dim, x, d, seed, init.data());
  }
  if (timings) timings[2] = secs(t0);
  if (init_out) std::memcpy(init_out, init.data(), n * dim * sizeof(float));

  t0 = clk::now();
  std::vector<float> emb(n * dim);
  orc_embedding(indptr.data(), ind.data(), wts.data(), n, dim, p->n_epochs, p->noise_level, 1.0f,
                0.1f, 0.5f, 0.9f, 1e-7f, init.data(), seed, emb.data());
  if (timings) timings[3] = secs(t0);
  if (embedding) std::memcpy(embedding, emb.data(), n * dim * sizeof(float));

  t0 = clk::now();
  int64_t nl;
  if (p->approx_n_clusters > 0) {
    orc_search_n_clusters(emb.data(), n, dim, p->min_samples, p->approx_n_clusters, labels,
                          strengths);
    persistence[0] = 0.0;
    nl = 1;
  } else {
    nl = orc_cluster_layers(emb.data(), n, dim, p->min_samples, p->base_min_cluster_size,
                            p->min_similarity_threshold, p->max_layers, p->persistence_fp64,
                            labels, strengths, persistence);
  }
  if (timings) timings[4] = secs(t0);
  return nl;
}
