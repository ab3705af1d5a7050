// Label-propagation initial layout (label_propagation_init, src/graph/label_prop.rs:537-564)
// for the plain form of evoc_hip. Every level of 64 or more nodes (label votes, outliers,
// partition, coarse graph, partition means, coarse-graph embedding, expansion) runs on the
// GPU, recursing level by level on device-resident graphs and data (top_level); the base
// case below 64 nodes (PCA or a random sphere) and a level whose coarse graph has no edges
// run here in host C++ with OpenMP (level). Accumulation orders follow the
// reference's sparse algebra (src/utils/sparse.rs) so the layout is reproducible:
//   * coarsening  R^T_norm G R   (per-partition dense accumulator, columns ascending),
//   * expansion   ((G o G^T) R_norm)_l1 L  and  (R_norm)_l1 L, averaged, centred, x0.1.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <unordered_map>

#include "host_stages.h"

namespace evoc {
namespace {

struct Sp {  // CSR with int32 columns
  std::vector<int64_t> ptr;
  std::vector<int32_t> col;
  std::vector<float> val;
  int64_t rows = 0, cols = 0;
};

// one Jacobi step of label propagation (label_prop.rs:45-86): unlabelled nodes take
// the heaviest neighbour label if its vote exceeds 1.0; ties go to the later label
std::vector<int64_t> lp_step(const Sp& g, const std::vector<int64_t>& lab) {
  std::vector<int64_t> out(g.rows);
#pragma omp parallel
  {
    std::vector<int64_t> vl;
    std::vector<float> vw;
#pragma omp for schedule(dynamic, 512)
    for (int64_t i = 0; i < g.rows; ++i) {
      if (lab[i] >= 0) {
        out[i] = lab[i];
        continue;
      }
      vl.clear();
      vw.clear();
      for (int64_t p = g.ptr[i]; p < g.ptr[i + 1]; ++p) {
        const int64_t l = lab[g.col[p]];
        size_t q = 0;
        while (q < vl.size() && vl[q] != l) ++q;
        if (q == vl.size()) {
          vl.push_back(l);
          vw.push_back(g.val[p]);
        } else {
          vw[q] += g.val[p];
        }
      }
      int64_t best = -1;
      float top = 1.0f;
      for (size_t q = 0; q < vl.size(); ++q) {
        if (vl[q] < 0) continue;
        if (vw[q] > top) {
          top = vw[q];
          best = vl[q];
        } else if (vw[q] == top) {
          best = vl[q];
        }
      }
      out[i] = best;
    }
  }
  return out;
}

// label_outliers (label_prop.rs:105-140), in node order, in place
void lp_outliers(const Sp& g, std::vector<int64_t>& lab, uint64_t seed) {
  int64_t maxl = -1;
  for (auto l : lab) maxl = std::max(maxl, l);
  if (maxl < 0) return;
  std::vector<int64_t> stack;
  for (int64_t i = 0; i < g.rows; ++i) {
    if (lab[i] >= 0) continue;
    stack.assign(1, i);
    bool found = false;
    for (int it = 0; it < 64 && !stack.empty() && !found; ++it) {
      const int64_t node = stack.back();
      stack.pop_back();
      for (int64_t p = g.ptr[node]; p < g.ptr[node + 1]; ++p) {
        const int64_t j = g.col[p];
        if (lab[j] >= 0) {
          lab[i] = lab[j];
          found = true;
          break;
        }
        stack.push_back(j);
      }
    }
    if (!found) {
      const uint32_t r = Philox::first(seed, (uint32_t)i, (uint32_t)(i >> 32), 0, TAG_LPOL);
      lab[i] = (int64_t)Philox::range(r, (uint64_t)maxl + 1);
    }
  }
}

// label_prop_loop (label_prop.rs:202-227): seeds, 20 steps, outliers, remap
std::vector<int32_t> lp_partition(const Sp& g, int64_t approx, int n_iter, uint64_t seed,
                                  int64_t* n_parts) {
  std::vector<int64_t> lab(g.rows, -1);
  for (int64_t t = 0; t < approx; ++t) {
    const uint32_t r = Philox::first(seed, (uint32_t)t, (uint32_t)(t >> 32), 0, TAG_LPSD);
    lab[Philox::range(r, (uint64_t)g.rows)] = t;
  }
  for (int it = 0; it < n_iter; ++it) lab = lp_step(g, lab);
  lp_outliers(g, lab, seed + 2000);
  // remap_labels (label_prop.rs:155-177): labelled ids in sorted order, -1s appended
  std::vector<int64_t> ids;
  ids.reserve(lab.size());
  for (auto l : lab)
    if (l >= 0) ids.push_back(l);
  std::sort(ids.begin(), ids.end());
  ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
  std::unordered_map<int64_t, int32_t> mp;
  mp.reserve(ids.size() * 2);
  for (size_t q = 0; q < ids.size(); ++q) mp[ids[q]] = (int32_t)q;
  int32_t next = (int32_t)ids.size();
  std::vector<int32_t> part(g.rows);
  for (int64_t i = 0; i < g.rows; ++i) part[i] = lab[i] < 0 ? next++ : mp[lab[i]];
  *n_parts = next;
  return part;
}

// members of each partition, ascending node order
void members(const std::vector<int32_t>& part, int64_t np, std::vector<int64_t>& ptr,
             std::vector<int32_t>& mem) {
  ptr.assign(np + 1, 0);
  for (auto a : part) ptr[a + 1]++;
  for (int64_t a = 0; a < np; ++a) ptr[a + 1] += ptr[a];
  mem.resize(part.size());
  std::vector<int64_t> cur(ptr.begin(), ptr.end() - 1);
  for (int64_t i = 0; i < (int64_t)part.size(); ++i) mem[cur[part[i]]++] = (int32_t)i;
}

// R^T_norm G R, clipped to [0, 1]
Sp coarsen(const Sp& g, const std::vector<int32_t>& part, int64_t np,
           const std::vector<int64_t>& mptr, const std::vector<int32_t>& mem,
           const std::vector<float>& s) {
  Sp r;
  r.rows = r.cols = np;
  std::vector<std::vector<std::pair<int32_t, float>>> rows(np);
#pragma omp parallel
  {
    std::vector<float> acc(g.cols, 0.0f);
    std::vector<char> hit(g.cols, 0);
    std::vector<int32_t> touched;
    std::vector<float> acc2(np, 0.0f);
    std::vector<char> hit2(np, 0);
    std::vector<int32_t> touched2;
#pragma omp for schedule(dynamic, 16)
    for (int64_t a = 0; a < np; ++a) {
      touched.clear();
      for (int64_t m = mptr[a]; m < mptr[a + 1]; ++m) {
        const int32_t i = mem[m];
        for (int64_t p = g.ptr[i]; p < g.ptr[i + 1]; ++p) {
          const int32_t j = g.col[p];
          const float pr = s[a] * g.val[p];
          if (!hit[j]) {
            hit[j] = 1;
            touched.push_back(j);
            acc[j] = pr;
          } else {
            acc[j] = acc[j] + pr;
          }
        }
      }
      std::sort(touched.begin(), touched.end());
      touched2.clear();
      for (int32_t j : touched) {
        const int32_t b = part[j];
        const float pr = acc[j] * 1.0f;
        if (!hit2[b]) {
          hit2[b] = 1;
          touched2.push_back(b);
          acc2[b] = pr;
        } else {
          acc2[b] = acc2[b] + pr;
        }
        hit[j] = 0;
        acc[j] = 0.0f;
      }
      std::sort(touched2.begin(), touched2.end());
      auto& row = rows[a];
      row.reserve(touched2.size());
      for (int32_t b : touched2) {
        float v = acc2[b];
        v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
        row.push_back({b, v});
        hit2[b] = 0;
        acc2[b] = 0.0f;
      }
    }
  }
  r.ptr.assign(np + 1, 0);
  for (int64_t a = 0; a < np; ++a) r.ptr[a + 1] = r.ptr[a] + (int64_t)rows[a].size();
  r.col.resize(r.ptr[np]);
  r.val.resize(r.ptr[np]);
  for (int64_t a = 0; a < np; ++a)
    for (size_t q = 0; q < rows[a].size(); ++q) {
      r.col[r.ptr[a] + q] = rows[a][q].first;
      r.val[r.ptr[a] + q] = rows[a][q].second;
    }
  return r;
}

// Gram-matrix PCA of a small centred f64 matrix (pca_init, label_prop.rs:288-321):
// eigenvectors by cyclic Jacobi, components ordered by eigenvalue, each component's
// sign fixed so its largest-magnitude entry is positive, then scaled to [-1, 1].
std::vector<float> pca(const std::vector<float>& x, int64_t n, int64_t d, int dim) {
  const int64_t kc = std::min<int64_t>(std::min<int64_t>(dim, n), d);
  std::vector<double> mu(d, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < d; ++j) mu[j] += (double)x[i * d + j];
  const double inv = 1.0 / (double)n;
  for (auto& m : mu) m *= inv;
  std::vector<double> a(n * n, 0.0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < n; ++j) {
      double acc = 0.0;
      for (int64_t t = 0; t < d; ++t)
        acc += ((double)x[i * d + t] - mu[t]) * ((double)x[j * d + t] - mu[t]);
      a[i * n + j] = acc;
    }
  std::vector<double> V(n * n, 0.0);
  for (int64_t i = 0; i < n; ++i) V[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int64_t p = 0; p < n; ++p)
      for (int64_t q = p + 1; q < n; ++q) off += a[p * n + q] * a[p * n + q];
    if (off < 1e-30) break;
    for (int64_t p = 0; p < n; ++p)
      for (int64_t q = p + 1; q < n; ++q) {
        const double apq = a[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double th = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (int64_t k = 0; k < n; ++k) {
          const double x1 = a[k * n + p], x2 = a[k * n + q];
          a[k * n + p] = c * x1 - sn * x2;
          a[k * n + q] = sn * x1 + c * x2;
        }
        for (int64_t k = 0; k < n; ++k) {
          const double x1 = a[p * n + k], x2 = a[q * n + k];
          a[p * n + k] = c * x1 - sn * x2;
          a[q * n + k] = sn * x1 + c * x2;
        }
        for (int64_t k = 0; k < n; ++k) {
          const double x1 = V[k * n + p], x2 = V[k * n + q];
          V[k * n + p] = c * x1 - sn * x2;
          V[k * n + q] = sn * x1 + c * x2;
        }
      }
  }
  std::vector<int64_t> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(),
                   [&](int64_t u, int64_t v) { return a[u * n + u] > a[v * n + v]; });
  std::vector<float> out(n * dim, 0.0f);
  for (int64_t c = 0; c < kc; ++c) {
    const int64_t col = ord[c];
    const double sv = std::sqrt(std::max(a[col * n + col], 0.0));
    int64_t arg = 0;
    for (int64_t i = 1; i < n; ++i)
      if (std::fabs(V[i * n + col]) > std::fabs(V[arg * n + col])) arg = i;
    const double sg = V[arg * n + col] < 0 ? -1.0 : 1.0;
    for (int64_t i = 0; i < n; ++i) out[i * dim + c] = (float)(sg * V[i * n + col] * sv);
    // normalise_to_unit_range (label_prop.rs:239-266)
    float lo = INFINITY, hi = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
      lo = std::fmin(lo, out[i * dim + c]);
      hi = std::fmax(hi, out[i * dim + c]);
    }
    const float range = hi - lo;
    if (range > 0.0f) {
      const float mid = (hi + lo) / 2.0f, half = range / 2.0f;
      for (int64_t i = 0; i < n; ++i) out[i * dim + c] = (out[i * dim + c] - mid) / half;
    }
  }
  return out;
}

// random_init (label_prop.rs:338-362): uniform [-1,1) then L2-normalised rows
std::vector<float> sphere_init(int64_t n, int dim, uint64_t seed) {
  std::vector<float> out(n * dim);
  std::vector<double> row(dim);
  for (int64_t i = 0; i < n; ++i) {
    double nrm = 0.0;
    for (int j = 0; j < dim; ++j) {
      row[j] = (double)(float)(Philox::unit(seed, (uint32_t)i, (uint32_t)j, 0, TAG_RINI) * 2.0 - 1.0);
      nrm += row[j] * row[j];
    }
    nrm = std::sqrt(nrm);
    for (int j = 0; j < dim; ++j) out[i * dim + j] = nrm > 0.0 ? (float)(row[j] / nrm) : (float)row[j];
  }
  return out;
}

// embedding of a coarse (possibly asymmetric) graph on the GPU: edges (i, j, w) with
// i < j taken from row i in row order (embedding.rs:79-90), one slot per endpoint
void coarse_embedding(const Sp& g, int dim, int epochs, float alpha, std::vector<float>& layout,
                      uint64_t seed, hipStream_t s, DevPool& pool) {
  const int64_t n = g.rows;
  std::vector<int64_t> deg(n, 0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t p = g.ptr[i]; p < g.ptr[i + 1]; ++p)
      if (i < g.col[p]) {
        deg[i]++;
        deg[g.col[p]]++;
      }
  std::vector<int64_t> ptr(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + deg[i];
  std::vector<int32_t> other(ptr[n]);
  std::vector<float> w(ptr[n]);
  std::vector<int64_t> cur(ptr.begin(), ptr.end() - 1);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t p = g.ptr[i]; p < g.ptr[i + 1]; ++p) {
      const int32_t j = g.col[p];
      if (i < j) {
        other[cur[i]] = j;
        w[cur[i]++] = g.val[p];
        other[cur[j]] = (int32_t)i;
        w[cur[j]++] = g.val[p];
      }
    }
  Graph dg;
  dg.n = n;
  dg.nnz = ptr[n];
  dg.indptr.alloc(pool, n + 1);
  dg.indices.alloc(pool, dg.nnz);
  dg.weights.alloc(pool, dg.nnz);
  Buf<float> de(pool, n * dim);
  EVOC_HIP(hipMemcpyAsync(dg.indptr.p, ptr.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
  if (dg.nnz) {
    EVOC_HIP(hipMemcpyAsync(dg.indices.p, other.data(), dg.nnz * sizeof(int32_t), hipMemcpyHostToDevice, s));
    EVOC_HIP(hipMemcpyAsync(dg.weights.p, w.data(), dg.nnz * sizeof(float), hipMemcpyHostToDevice, s));
  }
  EVOC_HIP(hipMemcpyAsync(de.p, layout.data(), n * dim * sizeof(float), hipMemcpyHostToDevice, s));
  EmbedParams ep;
  ep.n_epochs = epochs;
  ep.initial_alpha = alpha;
  ep.wide_deg = 1;  // a few thousand nodes: one wavefront each keeps their gathers in parallel
  embed_device(dg, dim, ep, de.p, false, seed, s, pool);
  EVOC_HIP(hipMemcpyAsync(layout.data(), de.p, n * dim * sizeof(float), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
}

// the same with the coarse graph in device memory: the slots are built there
void coarse_embedding_dev(const Graph& g, int dim, int epochs, float alpha,
                          std::vector<float>& layout, uint64_t seed, hipStream_t s, DevPool& pool) {
  const int64_t n = g.n;
  Graph dg;
  coarse_slots_device(g, dg, s, pool);
  Buf<float> de(pool, n * dim);
  EVOC_HIP(hipMemcpyAsync(de.p, layout.data(), n * dim * sizeof(float), hipMemcpyHostToDevice, s));
  EmbedParams ep;
  ep.n_epochs = epochs;
  ep.initial_alpha = alpha;
  ep.wide_deg = 1;  // a few thousand nodes: one wavefront each keeps their gathers in parallel
  embed_device(dg, dim, ep, de.p, false, seed, s, pool);
  EVOC_HIP(hipMemcpyAsync(layout.data(), de.p, n * dim * sizeof(float), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
}

// The data of a level: host rows (coarse levels) or device rows (the input level).
struct Rows {
  const std::vector<float>* host = nullptr;
  const float* dev = nullptr;
  int64_t ld = 0;
  bool present() const { return host || dev; }
};

// One level of label_propagation_init_inner on a host graph (the coarse levels). With
// dev (the same graph in device memory) and short rows the expansion runs on the GPU.
std::vector<float> level(const Sp& g, int dim, const Rows& data, int64_t d, int epochs,
                         int64_t approx, uint64_t seed, hipStream_t s, DevPool& pool,
                         int verbose, const Graph* dev = nullptr) {
  const int64_t n = g.rows;
  if (n < 64) {
    if (data.present() && dim <= d) {
      if (data.host && (int64_t)data.host->size() == n * d) return pca(*data.host, n, d, dim);
      if (data.dev) {
        std::vector<float> h(n * d);
        EVOC_HIP(hipMemcpy2DAsync(h.data(), d * sizeof(float), data.dev, data.ld * sizeof(float),
                                  d * sizeof(float), n, hipMemcpyDeviceToHost, s));
        EVOC_HIP(hipStreamSynchronize(s));
        return pca(h, n, d, dim);
      }
    }
    return sphere_init(n, dim, seed);
  }
  using tclk = std::chrono::steady_clock;
  auto tick = tclk::now();
  auto lap = [&](const char* what) {
    if (verbose >= 2) {
      const auto now = tclk::now();
      std::printf("    [n=%lld] %s: %.1f ms\n", (long long)n, what,
                  std::chrono::duration<double, std::milli>(now - tick).count());
      tick = now;
    }
  };
  int64_t np = 0;
  const std::vector<int32_t> part = lp_partition(g, approx, 20, seed, &np);
  lap("label propagation");
  if (verbose) std::printf("  Label prop: %lld nodes -> %lld partitions\n", (long long)n, (long long)np);
  std::vector<int64_t> mptr;
  std::vector<int32_t> mem;
  members(part, np, mptr, mem);
  // column L2 norm of the 0/1 partition map: 1/sqrt(|a|) (sparse.rs:437-469)
  std::vector<float> sa(np);
  for (int64_t a = 0; a < np; ++a) {
    float sq = 0.0f;
    for (int64_t m = mptr[a]; m < mptr[a + 1]; ++m) sq += 1.0f;
    const float nrm = std::sqrt(sq);
    sa[a] = nrm > 0.0f ? 1.0f / nrm : 1.0f;
  }
  const Sp coarse = coarsen(g, part, np, mptr, mem, sa);
  lap("coarsen graph");
  std::vector<float> cdata;
  if (data.present()) {
    // rows of R^T_norm, L1-normalised (sparse.rs:479-501), times the data
    std::vector<float> coef(np);
    for (int64_t a = 0; a < np; ++a) {
      float nrm = 0.0f;
      for (int64_t m = mptr[a]; m < mptr[a + 1]; ++m) nrm += std::fabs(sa[a]);
      coef[a] = nrm > 0.0f ? sa[a] * (1.0f / nrm) : sa[a];
    }
    cdata.assign(np * d, 0.0f);
    if (data.dev) {
      partition_means_device(data.dev, data.ld, d, mptr.data(), mem.data(), coef.data(), np, n,
                             cdata.data(), s, pool);
    } else {
#pragma omp parallel for schedule(dynamic, 8)
      for (int64_t a = 0; a < np; ++a) {
        float* out = &cdata[a * d];
        for (int64_t m = mptr[a]; m < mptr[a + 1]; ++m) {
          const float* x = &(*data.host)[(int64_t)mem[m] * d];
          for (int64_t c = 0; c < d; ++c) out[c] = out[c] + coef[a] * x[c];
        }
      }
    }
  }
  lap("coarsen data");
  const int ep = std::min(epochs, 255);
  Rows cr;
  if (data.present()) cr.host = &cdata;
  std::vector<float> layout = level(coarse, dim, cr, d, ep, approx / 4, seed + 3000, s, pool,
                                    verbose);
  tick = tclk::now();
  coarse_embedding(coarse, dim, ep, (float)(0.001 * (double)epochs), layout, seed + 4000, s, pool);
  lap("coarse embedding");
  // expansion: rows of (G o G^T) R_norm, L1-normalised, times the coarse layout,
  // averaged with the layout of each node's own partition
  std::vector<float> res(n * dim);
  // the GPU expansion compares a row's slots pairwise: it suits short rows only
  int64_t maxdeg = 0;
  for (int64_t i = 0; i < n; ++i) maxdeg = std::max<int64_t>(maxdeg, g.ptr[i + 1] - g.ptr[i]);
  if (dev && maxdeg <= 64) {
    Buf<int32_t> dpart(pool, n);
    Buf<float> dsa(pool, np), dres(pool, n * dim);
    EVOC_HIP(hipMemcpyAsync(dpart.p, part.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s));
    EVOC_HIP(hipMemcpyAsync(dsa.p, sa.data(), np * sizeof(float), hipMemcpyHostToDevice, s));
    expansion_device(dev->indptr.p, dev->indices.p, dev->weights.p, n, dev->nnz, dim, dpart.p,
                     dsa.p, np, layout.data(), dres.p, s, pool);
    EVOC_HIP(hipMemcpyAsync(res.data(), dres.p, n * dim * sizeof(float), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipStreamSynchronize(s));
    lap("expansion");
    return res;
  }
  // mate[p]: the position of G(j, i) for the entry p = (i, j), or -1. The transpose
  // (entries by column, rows ascending within one: per-thread counts and cursors) is
  // merged with each row, both in ascending order: O(nnz) in all.
  const int64_t nnz = g.ptr[n];
  std::vector<int64_t> mate(nnz, -1);
  {
    const int T = std::max(1, omp_get_max_threads());
    const size_t W = (size_t)n + 1;
    std::vector<int64_t> cur((size_t)T * W, 0);
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      int64_t* c = &cur[(size_t)t * W];
      for (int64_t i = n * t / T; i < n * (t + 1) / T; ++i)
        for (int64_t p = g.ptr[i]; p < g.ptr[i + 1]; ++p) ++c[g.col[p]];
    }
    std::vector<int64_t> tptr(n + 1);
    int64_t run = 0;
    for (int64_t j = 0; j < n; ++j) {
      tptr[j] = run;
      for (int t = 0; t < T; ++t) {
        const int64_t k = cur[(size_t)t * W + j];
        cur[(size_t)t * W + j] = run;
        run += k;
      }
    }
    tptr[n] = run;
    std::vector<int32_t> trow(std::max<int64_t>(nnz, 1));
    std::vector<int64_t> tpos(std::max<int64_t>(nnz, 1));
#pragma omp parallel num_threads(T)
    {
      const int t = omp_get_thread_num();
      int64_t* c = &cur[(size_t)t * W];
      for (int64_t i = n * t / T; i < n * (t + 1) / T; ++i)
        for (int64_t p = g.ptr[i]; p < g.ptr[i + 1]; ++p) {
          const int64_t q = c[g.col[p]]++;
          trow[q] = (int32_t)i;
          tpos[q] = p;
        }
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t j = 0; j < n; ++j) {
      int64_t a = g.ptr[j], b = tptr[j];
      const int64_t ae = g.ptr[j + 1], be = tptr[j + 1];
      while (a < ae && b < be) {
        const int32_t c = g.col[a], r = trow[b];
        if (c < r) {
          ++a;
        } else if (r < c) {
          ++b;
        } else {
          mate[tpos[b]] = a;  // entry (r, j) meets G(j, r)
          ++a;
          ++b;
        }
      }
    }
  }
#pragma omp parallel
  {
    std::vector<float> acc(np, 0.0f);
    std::vector<char> hit(np, 0);
    std::vector<int32_t> touched;
    std::vector<float> ex(dim), dr(dim);
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
      touched.clear();
      for (int64_t p = g.ptr[i]; p < g.ptr[i + 1]; ++p) {
        const int32_t j = g.col[p];
        // G^T(i, j) = G(j, i)
        const int64_t q = mate[p];
        if (q < 0) continue;
        const float sym = g.val[p] * g.val[q];
        const int32_t a = part[j];
        const float pr = sym * sa[a];
        if (!hit[a]) {
          hit[a] = 1;
          touched.push_back(a);
          acc[a] = pr;
        } else {
          acc[a] = acc[a] + pr;
        }
      }
      std::sort(touched.begin(), touched.end());
      float l1 = 0.0f;
      for (int32_t a : touched) l1 += std::fabs(acc[a]);
      const float inv = l1 > 0.0f ? 1.0f / l1 : 1.0f;
      for (int c = 0; c < dim; ++c) ex[c] = 0.0f;
      for (int32_t a : touched) {
        const float v = l1 > 0.0f ? acc[a] * inv : acc[a];
        for (int c = 0; c < dim; ++c) ex[c] = ex[c] + v * layout[(int64_t)a * dim + c];
        hit[a] = 0;
        acc[a] = 0.0f;
      }
      const int32_t pi = part[i];
      const float sv = sa[pi];
      const float dv = sv > 0.0f ? sv * (1.0f / std::fabs(sv)) : sv;
      for (int c = 0; c < dim; ++c) {
        dr[c] = 0.0f + dv * layout[(int64_t)pi * dim + c];
        res[i * dim + c] = (ex[c] + dr[c]) * 0.5f;
      }
    }
  }
  const float inv_n = (float)(1.0 / (double)n);
  for (int c = 0; c < dim; ++c) {
    float sum = 0.0f;
    for (int64_t i = 0; i < n; ++i) sum += res[i * dim + c];
    const float mean = sum * inv_n;
    for (int64_t i = 0; i < n; ++i) res[i * dim + c] = res[i * dim + c] - mean;
  }
  for (auto& v : res) v = 0.1f * v;
  lap("expansion");
  return res;
}

// One level with its graph in device memory: partition, coarse graph, partition means,
// coarse embedding and expansion on the GPU, recursing on the device through every coarse
// level of at least 64 nodes; below that the base case runs through level().
void top_level(const int64_t* d_ptr, const int32_t* d_col, const float* d_val, int64_t n,
               int64_t nnz, int dim, const Rows& data, int64_t d, int epochs, int64_t approx,
               uint64_t seed, float* d_out, hipStream_t s, DevPool& pool, int verbose) {
  if (n < 64) {  // the base case needs no graph
    Sp g;
    g.rows = g.cols = n;
    std::vector<float> r = level(g, dim, data, d, epochs, approx, seed, s, pool, verbose);
    EVOC_HIP(hipMemcpyAsync(d_out, r.data(), n * dim * sizeof(float), hipMemcpyHostToDevice, s));
    EVOC_HIP(hipStreamSynchronize(s));
    return;
  }
  using tclk = std::chrono::steady_clock;
  auto tick = tclk::now();
  auto lap = [&](const char* what) {
    if (verbose >= 2) {
      const auto now = tclk::now();
      std::printf("    [n=%lld] %s: %.1f ms\n", (long long)n, what,
                  std::chrono::duration<double, std::milli>(now - tick).count());
      tick = now;
    }
  };
  LpTop P;
  lp_top_partition(d_ptr, d_col, d_val, n, approx, 20, seed, P, s, pool);
  EVOC_HIP(hipStreamSynchronize(s));
  const int64_t np = P.np;
  lap("label propagation");
  if (verbose) std::printf("  Label prop: %lld nodes -> %lld partitions\n", (long long)n, (long long)np);
  Sp coarse;
  coarse.rows = coarse.cols = np;
  Graph dcoarse;  // the coarse graph stays in device memory for its expansion and embedding
  // a coarse level of 64 or more nodes recurses on the device and needs no host copy
  lp_top_coarsen(d_ptr, d_col, d_val, n, nnz, P, coarse.ptr, coarse.col, coarse.val, s, pool,
                 &dcoarse, np < 64);
  lap("coarsen graph");
  const Graph* dv = dcoarse.n == np ? &dcoarse : nullptr;  // not kept for an empty graph
  // the coarse level recurses on the device down to the base case (fewer than 64 nodes,
  // handled by level()); its graph and data stay in device memory
  const bool recurse = dv != nullptr && np >= 64;
  std::vector<float> cdata;
  Buf<float> dm;
  if (data.present()) {
    dm.alloc(pool, np * d);
    partition_means_dev(data.dev, data.ld, d, P.mptr.p, P.mem.p, P.coef.p, np, dm.p, s);
    if (!recurse) {
      cdata.resize(np * d);
      EVOC_HIP(hipMemcpyAsync(cdata.data(), dm.p, np * d * sizeof(float), hipMemcpyDeviceToHost,
                              s));
    }
    EVOC_HIP(hipStreamSynchronize(s));
  }
  lap("coarsen data");
  const int ep = std::min(epochs, 255);
  std::vector<float> layout;
  if (recurse) {
    Rows cr;
    if (data.present()) {
      cr.dev = dm.p;
      cr.ld = d;
    }
    Buf<float> dl(pool, np * dim);
    top_level(dv->indptr.p, dv->indices.p, dv->weights.p, np, dv->nnz, dim, cr, d, ep,
              approx / 4, seed + 3000, dl.p, s, pool, verbose);
    layout.resize(np * dim);
    EVOC_HIP(hipMemcpyAsync(layout.data(), dl.p, np * dim * sizeof(float), hipMemcpyDeviceToHost,
                            s));
    EVOC_HIP(hipStreamSynchronize(s));
  } else {
    Rows cr;
    if (data.present()) cr.host = &cdata;
    layout = level(coarse, dim, cr, d, ep, approx / 4, seed + 3000, s, pool, verbose, dv);
  }
  tick = tclk::now();
  if (dv)
    coarse_embedding_dev(*dv, dim, ep, (float)(0.001 * (double)epochs), layout, seed + 4000, s,
                         pool);
  else
    coarse_embedding(coarse, dim, ep, (float)(0.001 * (double)epochs), layout, seed + 4000, s,
                     pool);
  lap("coarse embedding");
  expansion_device(d_ptr, d_col, d_val, n, nnz, dim, P.part.p, P.sa.p, np, layout.data(), d_out,
                   s, pool);
  lap("expansion");
}

}  // namespace

void label_prop_init_device(const int64_t* d_indptr, const int32_t* d_indices,
                            const float* d_weights, int64_t nnz, int64_t n, int dim,
                            const float* data_dev, int64_t d, int64_t ld, uint64_t seed,
                            float* d_out, hipStream_t s, DevPool& pool, int verbose) {
  int64_t approx = (int64_t)(8.0 * std::sqrt((double)n));
  approx = std::min<int64_t>(std::max<int64_t>(approx, 256), 16384);
  approx = std::min<int64_t>(approx, n);
  Rows top;
  if (data_dev) {
    top.dev = data_dev;
    top.ld = ld;
  }
  top_level(d_indptr, d_indices, d_weights, n, nnz, dim, top, d, 50, approx, seed, d_out, s, pool,
            verbose);
}

void label_prop_init_host(const int64_t* indptr, const int32_t* indices, const float* weights,
                          int64_t n, int dim, const float* data_dev, int64_t d, int64_t ld,
                          uint64_t seed, float* out, hipStream_t s, DevPool& pool, int verbose) {
  const int64_t nnz = indptr[n];
  Buf<int64_t> dp(pool, n + 1);
  Buf<int32_t> dc(pool, std::max<int64_t>(nnz, 1));
  Buf<float> dw(pool, std::max<int64_t>(nnz, 1)), dout(pool, n * dim);
  EVOC_HIP(hipMemcpyAsync(dp.p, indptr, (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
  if (nnz) {
    EVOC_HIP(hipMemcpyAsync(dc.p, indices, nnz * sizeof(int32_t), hipMemcpyHostToDevice, s));
    EVOC_HIP(hipMemcpyAsync(dw.p, weights, nnz * sizeof(float), hipMemcpyHostToDevice, s));
  }
  label_prop_init_device(dp.p, dc.p, dw.p, nnz, n, dim, data_dev, d, ld, seed, dout.p, s, pool,
                         verbose);
  EVOC_HIP(hipMemcpyAsync(out, dout.p, n * dim * sizeof(float), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
}

}  // namespace evoc
