// The library's host tree code (evoc-rs_amd/csrc/hierarchy.cpp: linkage, the cluster-level
// steps over a condensation's splits, the approx_n_clusters bracket, and the full-row
// readers) built with AddressSanitizer and UndefinedBehaviorSanitizer and run on MSTs from
// the CPU oracle. No device is touched: the condensation's output (split records and
// every point's parent cluster and lambda) is taken from the oracle's condensed rows, and
// the device's label kernel is restated below as a loop; the layers and the search must
// then equal the oracle's (labels, strengths and fp64 persistence bit for bit).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../evoc-rs_amd/csrc/host_stages.h"
#include "../../oracle/evoc_oracle.h"

#define REQUIRE(c)                                                      \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      std::exit(1);                                                     \
    }                                                                   \
  } while (0)

static std::vector<float> embedding(int64_t n, int dim, int centres, uint32_t seed) {
  std::mt19937 g(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  std::vector<float> c(centres * dim), e(n * dim);
  for (auto& v : c) v = 8.0f * nd(g);
  for (int64_t i = 0; i < n; ++i)
    for (int j = 0; j < dim; ++j)
      e[i * dim + j] = c[(i % centres) * dim + j] + (0.3f + 0.2f * (i % 3)) * nd(g);
  // exact duplicates: equal weights in the MST and the linkage
  for (int64_t i = 0; i + 1 < n; i += 97) std::memcpy(&e[(i + 1) * dim], &e[i * dim], dim * 4);
  return e;
}

// the oracle's condensed rows at min_cluster_size m
static evoc::CTree oracle_rows(const evoc::Linkage& L, int64_t m) {
  const int64_t n = L.n, nr = (int64_t)L.left.size();
  evoc::CTree t;
  t.n = n;
  const int64_t cap = std::max<int64_t>(2 * n, 1);
  t.parent.resize(cap);
  t.child.resize(cap);
  t.size.resize(cap);
  t.lambda.resize(cap);
  const int64_t nn = nr == 0 ? 0
                             : orc_condense(L.left.data(), L.right.data(), L.dist.data(),
                                            L.size.data(), nr, n, m, t.parent.data(),
                                            t.child.data(), t.lambda.data(), t.size.data());
  t.parent.resize(nn);
  t.child.resize(nn);
  t.size.resize(nn);
  t.lambda.resize(nn);
  return t;
}

// what the device condensation returns (m >= 2): split records in row order, and per
// point its parent cluster and lambda
static evoc::DevCondensed as_condensed(const evoc::CTree& t, int64_t m, std::vector<int32_t>& pcl,
                                      std::vector<float>& plam) {
  evoc::DevCondensed c;
  c.n = t.n;
  c.mcs = m;
  c.E = (int64_t)t.parent.size();
  pcl.assign(t.n, -1);
  plam.assign(t.n, 0.0f);
  for (size_t i = 0; i < t.parent.size(); ++i) {
    if (t.size[i] > 1) {
      if (t.child[i] == t.n + 1 + 2 * c.S) {  // the left row of the next split
        REQUIRE(i + 1 < t.parent.size() && t.child[i + 1] == t.child[i] + 1);
        c.split_parent.push_back((int32_t)t.parent[i]);
        c.split_lam.push_back(t.lambda[i]);
        c.split_ls.push_back((int32_t)t.size[i]);
        c.split_rs.push_back((int32_t)t.size[i + 1]);
        ++c.S;
      }
    } else {
      REQUIRE(t.child[i] < t.n);
      pcl[t.child[i]] = (int32_t)t.parent[i];
      plam[t.child[i]] = t.lambda[i];
    }
  }
  for (int64_t p = 0; p < t.n && t.n >= 2; ++p) REQUIRE(pcl[p] >= t.n);
  return c;
}

// the device label kernel (tree_death_kernel + tree_labels_kernel) as a loop
static void label_loop(const evoc::DevCondensed& c, const std::vector<int32_t>& pcl,
                       const std::vector<float>& plam, int nl, const std::vector<int32_t>& clab,
                       const std::vector<int32_t>& dsel, const std::vector<int64_t>& nsel,
                       std::vector<int64_t>& lab, std::vector<float>& str,
                       std::vector<int64_t>& ncl, std::vector<int64_t>& assigned) {
  const int64_t n = c.n, C = 2 * c.S + 1;
  lab.assign(nl * n, -1);
  str.assign(nl * n, 0.0f);
  ncl.assign(nl, 0);
  assigned.assign(nl, 0);
  for (int l = 0; l < nl; ++l) {
    std::vector<float> death(nsel[l], 0.0f);
    for (int64_t p = 0; p < n; ++p) {
      const int32_t r = dsel[l * C + pcl[p] - n];
      if (r >= 0 && plam[p] > death[r]) death[r] = plam[p];
    }
    int64_t mx = -1;
    for (int64_t p = 0; p < n; ++p) {
      const int32_t lb = clab[l * C + pcl[p] - n];
      lab[l * n + p] = lb;
      if (lb >= 0) {
        const float ml = death[lb], lam = plam[p];
        str[l * n + p] = (ml == 0.0f || !std::isfinite(lam)) ? 1.0f : (lam < ml ? lam : ml) / ml;
        ++assigned[l];
        mx = std::max<int64_t>(mx, lb);
      }
    }
    ncl[l] = mx + 1;
  }
}

static void run_case(int64_t n, int dim, int centres, uint32_t seed) {
  const int ms = 5, mcs = 5, max_layers = 10;
  std::vector<float> e = embedding(n, dim, centres, seed);
  std::vector<int32_t> u(n), v(n);
  std::vector<float> w(n);
  // the kd-tree Boruvka MST (the same tree as dense Prim) for the larger cases
  orc_set_mst_algo(n > 5000 ? 1 : 0);
  const int64_t ne = orc_mst(e.data(), n, dim, ms, u.data(), v.data(), w.data());
  REQUIRE(ne == n - 1);
  const evoc::Linkage L = evoc::linkage_from_sorted(u.data(), v.data(), w.data(), ne, n);

  std::vector<int64_t> ol(max_layers * n);
  std::vector<float> os(max_layers * n);
  std::vector<double> op(max_layers);
  const int64_t onl = orc_cluster_layers(e.data(), n, dim, ms, mcs, 0.2, max_layers, 1,
                                         ol.data(), os.data(), op.data());
  evoc::Layers lay;
  if (n >= 2) {
    std::vector<int32_t> pcl;
    std::vector<float> plam;
    const evoc::CTree full = oracle_rows(L, mcs);
    const evoc::DevCondensed c = as_condensed(full, mcs, pcl, plam);
    // the split-based leaves are extract_leaves on the full rows
    REQUIRE(evoc::split_leaves(c) == evoc::leaves(full));
    const evoc::LayerSelection sel =
        evoc::select_layers(evoc::split_rows(c), evoc::split_leaves(c), mcs, 0.2, max_layers);
    std::vector<int32_t> clab, dsel;
    std::vector<int64_t> lab, nsel, ncl, assigned;
    std::vector<float> str;
    evoc::layer_tables(c, sel.sel, clab, dsel, nsel);
    label_loop(c, pcl, plam, (int)sel.sel.size(), clab, dsel, nsel, lab, str, ncl, assigned);
    lay = evoc::order_layers(n, lab.data(), str.data(), ncl, sel.pers);
    // the full-row readers (min_cluster_size < 2 path) agree on the same rows
    const evoc::Layers lr = evoc::layers_from_rows(full, evoc::select_layers(
                                                             full, evoc::leaves(full), mcs, 0.2,
                                                             max_layers));
    REQUIRE(lr.nl() == lay.nl() && lr.persistence == lay.persistence);
    REQUIRE(std::memcmp(lr.labels, lay.labels, lay.nl() * n * 8) == 0);
    REQUIRE(std::memcmp(lr.strengths, lay.strengths, lay.nl() * n * 4) == 0);
  } else {
    lay = evoc::Layers(n, 1);
    std::fill(lay.labels, lay.labels + n, -1);
    std::fill(lay.strengths, lay.strengths + n, 0.0f);
  }
  REQUIRE((int64_t)lay.nl() == onl);
  for (int64_t l = 0; l < onl; ++l) {
    REQUIRE(std::memcmp(lay.layer(l), &ol[l * n], n * 8) == 0);
    REQUIRE(std::memcmp(lay.strengths + l * n, &os[l * n], n * 4) == 0);
    REQUIRE(std::memcmp(&lay.persistence[l], &op[l], 8) == 0);
  }
  const int best = evoc::best_layer_index(lay.persistence);
  REQUIRE(best >= 0 && best < (int)onl);
  REQUIRE(evoc::adjusted_rand_index(lay.layer(0), lay.layer(0), n) == 1.0);

  // the leaf-count identity behind search_bounds, at every probe size
  for (int64_t m = 2; m <= std::max<int64_t>(n / 2, 2) && n >= 2; m += 1 + m / 8) {
    const evoc::CTree t = oracle_rows(L, m);
    int64_t splits = 0;
    for (size_t i = 0; i < L.left.size(); ++i) {
      const int64_t ls = L.left[i] >= n ? L.size[L.left[i] - n] : 1;
      const int64_t rs = L.right[i] >= n ? L.size[L.right[i] - n] : 1;
      splits += std::min(ls, rs) >= m;
    }
    REQUIRE((int64_t)evoc::leaves(t).size() == 1 + splits);
  }

  for (int64_t target : {2, 8, 40}) {
    int64_t lo = 2, hi = n / 2;
    evoc::search_bounds(L, target, lo, hi);
    int64_t nleaf[2] = {0, 0}, asg[2] = {0, 0};
    std::vector<int64_t> labs[2];
    std::vector<float> strs[2];
    const int64_t ms2[2] = {lo, hi};
    for (int q = 0; q < 2; ++q) {
      labs[q].assign(n, -1);
      strs[q].assign(n, 0.0f);
      if (n < 2) continue;
      const evoc::CTree t = oracle_rows(L, ms2[q]);
      const std::vector<int64_t> lv = evoc::leaves(t);
      nleaf[q] = (int64_t)lv.size();
      evoc::labels_and_strengths(t, lv, labs[q].data(), strs[q].data());
      for (auto x : labs[q]) asg[q] += x >= 0;
    }
    const int64_t dl = std::llabs(nleaf[0] - target), dh = std::llabs(nleaf[1] - target);
    const int pick = (dl < dh || (dl == dh && asg[0] >= asg[1])) ? 0 : 1;
    std::vector<int64_t> sl(n);
    std::vector<float> ss(n);
    REQUIRE(orc_search_n_clusters(e.data(), n, dim, ms, target, sl.data(), ss.data()) == 0);
    REQUIRE(std::memcmp(labs[pick].data(), sl.data(), n * 8) == 0);
    REQUIRE(std::memcmp(strs[pick].data(), ss.data(), n * 4) == 0);
  }
}

int main() {
  run_case(3000, 8, 12, 1);
  run_case(20000, 2, 400, 5);  // enough covered bins for the threaded persistence sums
  run_case(2500, 2, 30, 2);
  run_case(7, 3, 2, 3);      // fewer points than one min-cluster-size layer needs
  run_case(1, 4, 1, 4);      // a single point: no MST edge
  orc_set_mst_algo(0);
  std::vector<float> ok(64, 1.0f);
  evoc::check_finite_host(ok.data(), 64, "input");
  std::printf("host tree stages ran clean under the sanitizers and match the oracle\n");
  return 0;
}
