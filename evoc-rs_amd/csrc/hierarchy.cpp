// The host side of the cluster hierarchy: single linkage, and the cluster-level steps
// that read only the splits of a condensation (a few per thousand points): leaves,
// barcodes, persistence, peaks, the selection of diverse layers and the per-cluster label
// tables the device's label kernel reads. The condensation itself and the per-point labels
// and strengths run on the device (tree.hip, clusters.cpp); the readers of full condensed
// rows here serve min_cluster_size < 2, where every point is a cluster of its own.
//
//  linkage      src/clustering/linkage.rs:323-367   (union-find relabel, no parent walk)
//  leaves       src/clustering/condensed_tree.rs:264-294 (flag arrays)
//  labels       src/clustering/condensed_tree.rs:311-369 (nearest selected ancestor, the
//               representative the reference's union-find ends with)
//  strengths    src/clustering/condensed_tree.rs:384-484
//  barcodes     src/clustering/persistence.rs:42-118 (fp64)
//  persistence  src/clustering/persistence.rs:135-173 (fp64, barcode order per bin)
//  layers       src/clustering/persistence.rs:188-463
//  search       src/lib.rs:377-447 (leaf counts without condensing)
#include <omp.h>

#include <cstring>
#include <numeric>
#include <unordered_map>

#include "host_stages.h"

namespace evoc {

void linkage_from_sorted(const int32_t* u, const int32_t* v, const float* w, int64_t ne,
                         int64_t n, Linkage& L) {
  L.n = n;
  L.left.resize(ne);
  L.right.resize(ne);
  L.size.resize(ne);
  L.dist.resize(ne);
  // one record per vertex, so that a union-find step costs one cache line: the parent,
  // and at a root its rank, the linkage label of its component and the component size
  struct Vtx {
    int32_t par, rank;
    int64_t lab, size;
  };
  // kept by the calling thread between calls: a fresh allocation of this size is paid in
  // page faults on every call
  static thread_local std::vector<Vtx> V;
  V.resize(n);
  for (int64_t x = 0; x < n; ++x) V[x] = Vtx{(int32_t)x, 0, x, 1};
  auto find = [&](int64_t x) {
    while (V[x].par != x) {
      V[x].par = V[V[x].par].par;
      x = V[x].par;
    }
    return x;
  };
  constexpr int64_t AHEAD = 16;  // edges are in weight order: their vertices are scattered
  for (int64_t i = 0; i < ne; ++i) {
    if (i + AHEAD < ne) {
      __builtin_prefetch(&V[u[i + AHEAD]]);
      __builtin_prefetch(&V[v[i + AHEAD]]);
    }
    int64_t ru = find(u[i]), rv = find(v[i]);
    int64_t l = V[ru].lab, r = V[rv].lab;
    const int64_t ns = V[ru].size + V[rv].size;
    if (l < r) std::swap(l, r);
    L.left[i] = l;
    L.right[i] = r;
    L.dist[i] = w[i];
    L.size[i] = ns;
    if (V[ru].rank < V[rv].rank) std::swap(ru, rv);
    V[rv].par = (int32_t)ru;
    if (V[ru].rank == V[rv].rank) ++V[ru].rank;
    V[ru].lab = n + i;
    V[ru].size = ns;
  }
}

Linkage linkage_from_sorted(const int32_t* u, const int32_t* v, const float* w, int64_t ne,
                            int64_t n) {
  Linkage L;
  linkage_from_sorted(u, v, w, ne, n, L);
  return L;
}

// The clusters of a device condensation as condensed-tree rows, two per split in the
// walk's order: every row with more than one point, which is what barcodes() and the
// layer selection read.
CTree split_rows(const DevCondensed& c) {
  CTree t;
  t.n = c.n;
  const int64_t S = c.S;
  t.parent.resize(2 * S);
  t.child.resize(2 * S);
  t.lambda.resize(2 * S);
  t.size.resize(2 * S);
  for (int64_t s = 0; s < S; ++s) {
    t.parent[2 * s] = t.parent[2 * s + 1] = c.split_parent[s];
    t.child[2 * s] = c.n + 1 + 2 * s;
    t.child[2 * s + 1] = c.n + 2 + 2 * s;
    t.lambda[2 * s] = t.lambda[2 * s + 1] = c.split_lam[s];
    t.size[2 * s] = c.split_ls[s];
    t.size[2 * s + 1] = c.split_rs[s];
  }
  return t;
}

// extract_leaves for min_cluster_size >= 2: every cluster n .. n + 2S parents at least one
// row (the node it labels is processed and emits rows), and only splits emit rows of more
// than one point, so the leaves are the clusters that parent no split, in id order.
std::vector<int64_t> split_leaves(const DevCondensed& c) {
  std::vector<int64_t> out;
  if (c.n < 2) return out;  // no linkage rows, no condensed tree
  std::vector<char> isp(2 * c.S + 1, 0);
  for (int64_t s = 0; s < c.S; ++s) isp[c.split_parent[s] - c.n] = 1;
  for (int64_t q = 0; q <= 2 * c.S; ++q)
    if (!isp[q]) out.push_back(c.n + q);
  return out;
}

// Per layer and cluster: the label of the points that fall out of the cluster (the rank of
// its nearest selected ancestor-or-self among the layer's sorted selection, -1 when that is
// the root or there is none) and the cluster's own rank (-1 if not selected). Parents have
// smaller ids than their children, so one pass in id order resolves the ancestors.
void layer_tables(const DevCondensed& c, const std::vector<std::vector<int64_t>>& sel,
                  std::vector<int32_t>& clab, std::vector<int32_t>& dsel,
                  std::vector<int64_t>& nsel) {
  const int64_t n = c.n, C = 2 * c.S + 1, nl = (int64_t)sel.size();
  clab.assign(nl * C, -1);
  dsel.assign(nl * C, -1);
  nsel.assign(nl, 0);
  std::vector<int64_t> near(C);
  for (int64_t l = 0; l < nl; ++l) {
    std::vector<int64_t> sorted(sel[l]);
    std::sort(sorted.begin(), sorted.end());
    nsel[l] = (int64_t)sorted.size();
    int32_t* rk = dsel.data() + l * C;
    for (size_t q = 0; q < sorted.size(); ++q)
      if (sorted[q] >= n && sorted[q] < n + C) rk[sorted[q] - n] = (int32_t)q;
    for (int64_t q = 0; q < C; ++q) {
      if (rk[q] >= 0) near[q] = q;
      else if (q > 0) near[q] = near[c.split_parent[(q - 1) / 2] - n];
      else near[q] = -1;
    }
    int32_t* cl = clab.data() + l * C;
    for (int64_t q = 0; q < C; ++q) cl[q] = near[q] > 0 ? rk[near[q]] : -1;
  }
}

// leaves(condense(L, m)).size() for m >= 2 is 1 + the number of linkage rows whose smaller
// child holds at least m points: those rows are exactly the splits, and each split adds two
// clusters of which its parent stops being a leaf. So the approx_n_clusters binary search
// (src/lib.rs:377-447) needs no condensation per probe.
void search_bounds(const Linkage& L, int64_t target_k, int64_t& lo, int64_t& hi) {
  const int64_t n = L.n, nr = (int64_t)L.left.size();
  lo = 2;
  hi = n / 2;
  if (hi - lo <= 1) return;
  std::vector<int64_t> ge(hi + 2, 0);
  for (int64_t i = 0; i < nr; ++i) {
    const int64_t ls = L.left[i] >= n ? L.size[L.left[i] - n] : 1;
    const int64_t rs = L.right[i] >= n ? L.size[L.right[i] - n] : 1;
    ++ge[std::min(std::min(ls, rs), hi + 1)];
  }
  for (int64_t m = hi; m >= 0; --m) ge[m] += ge[m + 1];
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (mid == lo || mid == hi) break;
    const int64_t mk = 1 + ge[mid];
    if (mk < target_k) hi = mid;
    else lo = mid;
  }
}

std::vector<int64_t> leaves(const CTree& t) {
  std::vector<int64_t> out;
  if (t.parent.empty()) return out;
  int64_t max_id = 0;
  for (size_t i = 0; i < t.parent.size(); ++i) max_id = std::max({max_id, t.parent[i], t.child[i]});
  std::vector<char> is_parent(max_id + 1, 0), appears(max_id + 1, 0);
  for (size_t i = 0; i < t.parent.size(); ++i) {
    if (t.size[i] > 1) is_parent[t.parent[i]] = 1;
    appears[t.parent[i]] = 1;
  }
  for (int64_t c = t.n; c <= max_id; ++c)
    if (!is_parent[c] && appears[c]) out.push_back(c);
  return out;
}

void labels_and_strengths(const CTree& t, const std::vector<int64_t>& clusters, int64_t* labels,
                          float* strengths) {
  const int64_t n = t.n;
  std::fill(labels, labels + n, -1);
  if (strengths) std::fill(strengths, strengths + n, 0.0f);
  if (clusters.empty() || t.parent.empty()) return;
  const size_t nr = t.parent.size();
  int64_t max_id = 0, root = INT64_MAX;
  for (size_t i = 0; i < nr; ++i) {
    max_id = std::max({max_id, t.parent[i], t.child[i]});
    root = std::min(root, t.parent[i]);
  }
  std::vector<int64_t> sorted(clusters);
  std::sort(sorted.begin(), sorted.end());
  // per-cluster arrays indexed by id - n (cluster ids are n .. max_id)
  const int64_t nc = max_id + 1 - n;
  if (nc <= 0) return;
  std::vector<int64_t> rank(nc, -1);
  for (size_t q = 0; q < sorted.size(); ++q)
    if (sorted[q] >= n && sorted[q] <= max_id) rank[sorted[q] - n] = (int64_t)q;
  // nearest selected ancestor of every cluster id (parents precede children)
  std::vector<int64_t> cpar(nc, -1);
  for (size_t i = 0; i < nr; ++i)
    if (t.child[i] >= n) cpar[t.child[i] - n] = t.parent[i];
  std::vector<int64_t> sel(nc, -1);
  for (int64_t c = 0; c < nc; ++c) {
    if (rank[c] >= 0) sel[c] = c + n;
    else if (cpar[c] >= n) sel[c] = sel[cpar[c] - n];
  }
  for (size_t i = 0; i < nr; ++i) {
    const int64_t p = t.child[i];
    if (p >= n) continue;
    const int64_t s = sel[t.parent[i] - n];
    labels[p] = (s > root) ? rank[s - n] : -1;
  }
  if (!strengths) return;
  std::vector<float> death(nc, 0.0f);
  for (size_t i = 0; i < nr; ++i)
    if (t.size[i] == 1 && rank[t.parent[i] - n] >= 0 && t.lambda[i] > death[t.parent[i] - n])
      death[t.parent[i] - n] = t.lambda[i];
  for (size_t i = 0; i < nr; ++i) {
    const int64_t pt = t.child[i];
    if (pt >= root || pt >= n) continue;
    const int64_t lb = labels[pt];
    if (lb < 0) continue;
    const float ml = death[sorted[lb] - n];
    const float lam = t.lambda[i];
    if (ml == 0.0f || !std::isfinite(lam)) strengths[pt] = 1.0f;
    else strengths[pt] = (lam < ml ? lam : ml) / ml;
  }
}

namespace {

struct Bar {
  std::vector<double> birth, death, lam;
};

Bar barcodes(const CTree& t, int64_t min_size) {
  Bar b;
  std::vector<size_t> cn;
  for (size_t i = 0; i < t.parent.size(); ++i)
    if (t.size[i] > 1) cn.push_back(i);
  if (cn.empty()) return b;
  int64_t max_child = 0;
  for (auto i : cn) max_child = std::max(max_child, t.child[i]);
  const int64_t nn = max_child - t.n + 1;
  b.birth.assign(nn, (double)min_size);
  b.death.assign(nn, 0.0);
  b.lam.assign(nn, 0.0);
  b.death[0] = (double)t.n;
  int64_t i = (int64_t)cn.size();
  while (i >= 2) {
    i -= 2;
    const size_t a = cn[i], c = cn[i + 1];
    const int64_t ia = t.child[a] - t.n, ib = t.child[c] - t.n;
    const double lv = std::exp(-1.0 / (double)t.lambda[a]);
    b.lam[ia] = lv;
    b.lam[ib] = lv;
    const double dt = (double)std::min(t.size[a], t.size[c]);
    b.death[ia] = dt;
    b.death[ib] = dt;
    const int64_t pi = t.parent[a] - t.n;
    const double mb = std::max(std::max(b.birth[ia], b.birth[ib]), dt);
    b.birth[pi] = std::max(b.birth[pi], mb);
  }
  return b;
}

// Total persistence per size bin in fp64: a bin's value is the sum, in barcode order, of
// the contributions of the barcodes that cover it. Walking the barcodes in order and
// adding each contribution to the bins it covers performs exactly those adds, in that
// order, for every bin; threads own disjoint ranges of bins, so the order is kept.
void total_persistence(const Bar& b, std::vector<double>& sizes, std::vector<double>& pers) {
  sizes.assign(b.birth.begin(), b.birth.end());
  std::sort(sizes.begin(), sizes.end());
  sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
  std::vector<int64_t> bp, dp;
  std::vector<double> contrib;
  for (size_t i = 1; i < b.birth.size(); ++i) {
    if (b.death[i] <= b.birth[i]) continue;
    bp.push_back(std::lower_bound(sizes.begin(), sizes.end(), b.birth[i]) - sizes.begin());
    dp.push_back(std::lower_bound(sizes.begin(), sizes.end(), b.death[i]) - sizes.begin());
    contrib.push_back((b.death[i] - b.birth[i]) * b.lam[i]);
  }
  pers.assign(sizes.size(), 0.0);
  const int64_t nbins = (int64_t)sizes.size(), nbc = (int64_t)bp.size();
  int64_t cover = 0;
  for (int64_t q = 0; q < nbc; ++q) cover += dp[q] - bp[q];
#pragma omp parallel if (cover > (1 << 14))
  {
    const int64_t T = omp_get_num_threads(), t = omp_get_thread_num();
    const int64_t lo = nbins * t / T, hi = nbins * (t + 1) / T;
    double* out = pers.data();
    for (int64_t q = 0; q < nbc; ++q) {
      const int64_t a = std::max(bp[q], lo), e = std::min(dp[q], hi);
      const double c = contrib[q];
      for (int64_t k = a; k < e; ++k) out[k] = out[k] + c;
    }
  }
}

std::vector<int64_t> find_peaks(const std::vector<double>& x) {
  std::vector<int64_t> pk;
  if (x.size() < 3) return pk;
  size_t i = 1;
  const size_t imax = x.size() - 1;
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

double jaccard(const Bar& b, double x, double y) {
  int64_t inter = 0, uni = 0;
  for (size_t i = 0; i < b.birth.size(); ++i) {
    const bool ia = b.birth[i] <= x && b.death[i] > x;
    const bool ib = b.birth[i] <= y && b.death[i] > y;
    inter += ia && ib;
    uni += ia || ib;
  }
  return uni == 0 ? 0.0 : (double)inter / (double)uni;
}


}  // namespace

LayerSelection select_layers(const CTree& t, std::vector<int64_t> base, int64_t base_mcs,
                             double min_sim, int64_t max_layers) {
  const int64_t n = t.n;
  LayerSelection out;
  // the base layer (the leaves, persistence 0), then the diverse peaks of the
  // total-persistence curve, each layer its active clusters
  out.sel.push_back(std::move(base));
  out.pers.push_back(0.0);
  const Bar b = barcodes(t, base_mcs);
  if (b.birth.empty()) return out;
  std::vector<double> sizes, curve;
  total_persistence(b, sizes, curve);
  const std::vector<int64_t> pk = find_peaks(curve);
  std::vector<std::pair<int64_t, double>> idx;
  for (auto p : pk) idx.push_back({p, curve[p]});
  std::stable_sort(idx.begin(), idx.end(), [](const std::pair<int64_t, double>& a,
                                              const std::pair<int64_t, double>& c) {
    return c.second < a.second;
  });
  const int64_t cap = max_layers > 0 ? max_layers - 1 : 0;
  std::vector<double> chosen;
  for (auto& pp : idx) {
    if ((int64_t)chosen.size() >= cap) break;
    const double bs = sizes[pp.first];
    bool diverse = true;
    for (double c : chosen)
      if (!(jaccard(b, bs, c) <= min_sim)) {
        diverse = false;
        break;
      }
    if (!diverse) continue;
    chosen.push_back(bs);
    std::vector<int64_t> act;
    for (size_t i = 0; i < b.birth.size(); ++i)
      if (b.birth[i] <= bs && b.death[i] > bs) act.push_back((int64_t)i + n);
    out.sel.push_back(std::move(act));
    out.pers.push_back(curve[pp.first]);
  }
  return out;
}

void parallel_copy(void* dst, const void* src, size_t bytes) {
  constexpr size_t CH = size_t(1) << 21;
  const int64_t nch = (int64_t)((bytes + CH - 1) / CH);
#pragma omp parallel for schedule(static) if (nch > 1)
  for (int64_t c = 0; c < nch; ++c) {
    const size_t a = (size_t)c * CH, e = std::min(bytes, a + CH);
    std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, e - a);
  }
}

Layers::Layers(int64_t n_, int nl) : n(n_) {
  const size_t m = (size_t)std::max(nl, 1) * (size_t)std::max<int64_t>(n_, 1);
  labels = static_cast<int64_t*>(std::malloc(m * sizeof(int64_t)));
  strengths = static_cast<float*>(std::malloc(m * sizeof(float)));
  if (!labels || !strengths) {
    std::free(labels);
    std::free(strengths);
    throw std::bad_alloc();
  }
  persistence.assign(nl, 0.0);
}
Layers::~Layers() {
  std::free(labels);
  std::free(strengths);
}
Layers::Layers(Layers&& o) noexcept
    : n(o.n), labels(o.labels), strengths(o.strengths), persistence(std::move(o.persistence)) {
  o.labels = nullptr;
  o.strengths = nullptr;
}
Layers& Layers::operator=(Layers&& o) noexcept {
  if (this != &o) {
    std::free(labels);
    std::free(strengths);
    n = o.n;
    labels = o.labels;
    strengths = o.strengths;
    persistence = std::move(o.persistence);
    o.labels = nullptr;
    o.strengths = nullptr;
  }
  return *this;
}

Layers order_layers(int64_t n, const int64_t* lab, const float* str,
                    const std::vector<int64_t>& ncl, const std::vector<double>& pers) {
  const int nl = (int)ncl.size();
  const std::vector<int64_t> order = order_by_clusters(ncl);
  Layers out(n, nl);
  for (int q = 0; q < nl; ++q) {
    const int64_t o = order[q];
    std::memcpy(out.labels + (size_t)q * n, lab + (size_t)o * n, n * sizeof(int64_t));
    std::memcpy(out.strengths + (size_t)q * n, str + (size_t)o * n, n * sizeof(float));
    out.persistence[q] = pers[o];
  }
  return out;
}

// Labels and strengths of every selected layer from the full rows on the host (used where
// points are clusters of their own, min_cluster_size < 2, so no point falls out).
Layers layers_from_rows(const CTree& t, const LayerSelection& s) {
  const int64_t n = t.n, nl = (int64_t)s.sel.size();
  std::vector<int64_t> lab(nl * n);
  std::vector<float> str(nl * n);
  std::vector<int64_t> nc(nl);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t l = 0; l < nl; ++l) {
    labels_and_strengths(t, s.sel[l], lab.data() + l * n, str.data() + l * n);
    int64_t m = -1;
    for (int64_t i = 0; i < n; ++i) m = std::max(m, lab[l * n + i]);
    nc[l] = m + 1;
  }
  return order_layers(n, lab.data(), str.data(), nc, s.pers);
}

int best_layer_index(const std::vector<double>& p) {
  if (p.size() <= 1) return 0;
  int best = 0;
  for (size_t i = 1; i < p.size(); ++i)
    if (!(p[i] < p[best])) best = (int)i;  // Iterator::max_by keeps the last maximum
  return best;
}

// Adjusted Rand index (Hubert and Arabie, 1985) of two labellings; every label value,
// -1 included, is a cluster of its own. Two labellings with one cluster each score 1.
double adjusted_rand_index(const int64_t* a, const int64_t* b, int64_t n) {
  if (n <= 1) return 1.0;
  std::unordered_map<int64_t, int64_t> ca, cb;
  std::unordered_map<uint64_t, int64_t> cab;
  std::unordered_map<int64_t, int64_t> ia, ib;
  for (int64_t i = 0; i < n; ++i) {
    auto ra = ia.emplace(a[i], (int64_t)ia.size()).first->second;
    auto rb = ib.emplace(b[i], (int64_t)ib.size()).first->second;
    ca[ra]++;
    cb[rb]++;
    cab[((uint64_t)ra << 32) | (uint64_t)rb]++;
  }
  auto c2 = [](int64_t m) { return 0.5 * (double)m * (double)(m - 1); };
  double sab = 0, sa = 0, sb = 0;
  for (auto& kv : cab) sab += c2(kv.second);
  for (auto& kv : ca) sa += c2(kv.second);
  for (auto& kv : cb) sb += c2(kv.second);
  const double expected = sa * sb / c2(n);
  const double mx = 0.5 * (sa + sb);
  if (mx == expected) return 1.0;
  return (sab - expected) / (mx - expected);
}

}  // namespace evoc
