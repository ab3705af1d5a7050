// The library's host tree stages (evoc-rs_amd/csrc/hierarchy.cpp: linkage, condensed
// tree, leaves, labels, strengths, persistence, layers, the n-cluster search) built with
// AddressSanitizer and UndefinedBehaviorSanitizer and run on MSTs from the CPU oracle,
// with the results required to equal the oracle's (labels, strengths and fp64
// persistence bit for bit). No device is touched: these stages take host arrays.
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
  const evoc::Layers lay = evoc::cluster_layers_from_linkage(L, mcs, 0.2, max_layers);

  std::vector<int64_t> ol(max_layers * n);
  std::vector<float> os(max_layers * n);
  std::vector<double> op(max_layers);
  const int64_t onl = orc_cluster_layers(e.data(), n, dim, ms, mcs, 0.2, max_layers, 1,
                                         ol.data(), os.data(), op.data());
  REQUIRE((int64_t)lay.labels.size() == onl);
  for (int64_t l = 0; l < onl; ++l) {
    REQUIRE(std::memcmp(lay.labels[l].data(), &ol[l * n], n * 8) == 0);
    REQUIRE(std::memcmp(lay.strengths[l].data(), &os[l * n], n * 4) == 0);
    REQUIRE(std::memcmp(&lay.persistence[l], &op[l], 8) == 0);
  }
  const int best = evoc::best_layer_index(lay.persistence);
  REQUIRE(best >= 0 && best < (int)onl);

  // condensed tree readers directly
  const evoc::CTree t = evoc::condense(L, mcs);
  const std::vector<int64_t> lv = evoc::leaves(t);
  std::vector<int64_t> labels(n);
  std::vector<float> st(n);
  evoc::labels_and_strengths(t, lv, labels.data(), st.data());
  REQUIRE(evoc::adjusted_rand_index(labels.data(), labels.data(), n) == 1.0);

  for (int64_t target : {2, 8, 40}) {
    std::vector<int64_t> hl(n), sl(n);
    std::vector<float> hs(n), ss(n);
    evoc::search_n_clusters_from_linkage(L, target, hl.data(), hs.data());
    REQUIRE(orc_search_n_clusters(e.data(), n, dim, ms, target, sl.data(), ss.data()) == 0);
    REQUIRE(std::memcmp(hl.data(), sl.data(), n * 8) == 0);
    REQUIRE(std::memcmp(hs.data(), ss.data(), n * 4) == 0);
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
