// Runs every stage of the CPU oracle (oracle/evoc_oracle.cpp) on small seeded inputs,
// built with AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize/Makefile).
// A sanitizer report aborts the run; the checks here only make sure each call did its
// work (return codes, sizes), the results themselves are checked by the Python tests.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../oracle/evoc_oracle.h"

#define STEP(name) (std::fprintf(stderr, "oracle: %s\n", name), std::fflush(stderr))

#define REQUIRE(c)                                                      \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      std::exit(1);                                                     \
    }                                                                   \
  } while (0)

static std::vector<float> blobs(int64_t n, int64_t d, int centres, uint32_t seed) {
  std::mt19937 g(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  std::vector<float> c(centres * d), x(n * d);
  for (auto& v : c) v = 6.0f * nd(g);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < d; ++j) x[i * d + j] = c[(i % centres) * d + j] + nd(g);
  return x;
}

int main() {
  const int64_t n = 1200, d = 24;
  const int k = 15, dim = 4, ms = 5;
  std::vector<float> x = blobs(n, d, 8, 7);

  // kNN: exhaustive, sampled rows, NN-Descent (both metrics)
  STEP("knn");
  std::vector<int32_t> idx(n * k);
  std::vector<float> dist(n * k);
  for (int metric = 0; metric < 2; ++metric) {
    REQUIRE(orc_knn_exhaustive(x.data(), n, d, metric, k, idx.data(), dist.data()) == 0);
    std::vector<int64_t> rows = {0, 17, n - 1};
    std::vector<int32_t> ri(3 * k);
    std::vector<float> rd(3 * k);
    REQUIRE(orc_knn_rows(x.data(), n, d, metric, k, rows.data(), 3, ri.data(), rd.data()) == 0);
    std::vector<int32_t> ni(n * k);
    std::vector<float> ndist(n * k);
    int iters = 0;
    REQUIRE(orc_knn_nndescent(x.data(), n, d, metric, k, 42, 8, 0.001f, 1.0f, ni.data(),
                              ndist.data(), &iters) == 0);
  }
  // k larger than n - 1 pads with -1
  {
    std::vector<int32_t> si(5 * 8);
    std::vector<float> sd(5 * 8);
    REQUIRE(orc_knn_exhaustive(x.data(), 5, d, 0, 8, si.data(), sd.data()) == 0);
    REQUIRE(si[7] == -1);
  }
  REQUIRE(orc_knn_exhaustive(x.data(), n, d, 0, k, idx.data(), dist.data()) == 0);

  // fuzzy set, f32 and f64, symmetrised and directed
  STEP("fuzzy");
  std::vector<int64_t> ip(n + 1);
  std::vector<int32_t> ix(2 * n * k);
  std::vector<float> w(2 * n * k), sig(n), rho(n);
  int64_t nnz = 0;
  for (int sym = 0; sym < 2; ++sym) {
    nnz = orc_fuzzy_graph(idx.data(), dist.data(), n, k, (float)k, sym, sig.data(), rho.data(),
                          ip.data(), ix.data(), w.data());
    REQUIRE(nnz > 0);
  }
  {
    std::vector<double> d64(dist.begin(), dist.end()), s64(n), r64(n), w64(2 * n * k);
    std::vector<int64_t> ip64(n + 1);
    std::vector<int32_t> ix64(2 * n * k);
    REQUIRE(orc_fuzzy_graph_f64(idx.data(), d64.data(), n, k, (float)k, 1, s64.data(),
                                r64.data(), ip64.data(), ix64.data(), w64.data()) == nnz);
  }
  {
    std::vector<int64_t> rr = {0, 1, 2, 2};
    std::vector<int32_t> cc = {1, 0, 0, 3};
    std::vector<float> vv = {0.8f, 0.6f, 0.7f, 0.2f};
    std::vector<int64_t> sip(5);
    std::vector<int32_t> six(8);
    std::vector<float> sw(8);
    REQUIRE(orc_symmetrise(rr.data(), cc.data(), vv.data(), 4, 4, sip.data(), six.data(),
                           sw.data()) == 6);
  }

  // sparse algebra and label-propagation pieces on the fuzzy graph
  STEP("csr");
  std::vector<int64_t> ix64(ix.begin(), ix.begin() + nnz);
  for (int op = 0; op <= 5; ++op) {
    const int64_t cap = 8 * n * k * k;
    std::vector<int64_t> oip(n + 1), oix(cap), shape(2);
    std::vector<float> odv(cap);
    const int64_t r = orc_csr_op(op, ip.data(), ix64.data(), w.data(), n, n, ip.data(),
                                 ix64.data(), w.data(), n, n, oip.data(), oix.data(),
                                 odv.data(), cap, shape.data());
    REQUIRE(r >= 0);
  }
  std::vector<int64_t> lab(n, -1), lab2(n);
  STEP("label propagation");
  for (int64_t i = 0; i < n; i += 3) lab[i] = i % 5;
  REQUIRE(orc_lp_iteration(ip.data(), ix64.data(), w.data(), n, lab.data(), lab2.data()) == 0);
  REQUIRE(orc_lp_outliers(ip.data(), ix64.data(), w.data(), n, lab2.data(), 3) == 0);
  const int64_t np = orc_lp_remap(lab2.data(), n);
  REQUIRE(np > 0);
  std::vector<int64_t> part(n);
  const int64_t nparts = orc_lp_loop(ip.data(), ix64.data(), w.data(), n, 40, 20, 5, part.data());
  REQUIRE(nparts > 0);
  {
    std::vector<int64_t> pip(n + 1), pix(n);
    std::vector<float> pdv(n);
    REQUIRE(orc_csr_from_partition(part.data(), n, nparts, pip.data(), pix.data(),
                                   pdv.data()) == 0);
    std::vector<float> prod(n * d);   // (n x nparts) times the first nparts rows of x
    REQUIRE(orc_csr_matmul_dense(pip.data(), pix.data(), pdv.data(), n, nparts, x.data(), d,
                                 prod.data()) == 0);
    std::vector<float> lay(nparts * dim), out(n * dim);
    REQUIRE(orc_random_init(nparts, dim, 9, lay.data()) == 0);
    REQUIRE(orc_lp_expand(ip.data(), ix64.data(), w.data(), n, part.data(), nparts, lay.data(),
                          dim, 0.1, out.data()) == 0);
    std::vector<int64_t> rows = {0, 1, 1, 2}, cols = {1, 0, 2, 1};
    std::vector<float> vals = {1, 2, 3, 4};
    std::vector<int64_t> cip(4), cix(4);
    std::vector<float> cdv(4);
    REQUIRE(orc_csr_from_coo(rows.data(), cols.data(), vals.data(), 4, 3, cip.data(), cix.data(),
                             cdv.data()) == 4);
  }
  {
    // the PCA base case sees a few dozen coarse nodes (Jacobi on the Gram matrix)
    const int64_t m_rows = 80;
    std::vector<float> m(m_rows * dim);
    REQUIRE(orc_pca_init(x.data(), m_rows, d, dim, m.data()) == 0);
    REQUIRE(orc_unit_range(m.data(), m_rows, dim) == 0);
  }

  // initial layout and embedding
  STEP("layout");
  std::vector<float> init(n * dim), emb(n * dim);
  REQUIRE(orc_label_prop_init(ip.data(), ix.data(), w.data(), n, dim, x.data(), d, 42,
                              init.data()) == 0);
  REQUIRE(orc_embedding(ip.data(), ix.data(), w.data(), n, dim, 30, 0.5f, 1.0f, 0.1f, 0.5f,
                        0.9f, 1e-7f, init.data(), 42, emb.data()) == 0);

  // core distances, MSTs (Prim, Boruvka, f64), linkage, condensed tree and its readers
  STEP("mst");
  std::vector<float> core(n), core_kd(n);
  REQUIRE(orc_core_sq(emb.data(), n, dim, ms, core.data()) == 0);
  REQUIRE(orc_core_sq_kd(emb.data(), n, dim, ms, core_kd.data()) == 0);
  std::vector<int32_t> u(n), v(n);
  std::vector<float> wt(n);
  REQUIRE(orc_mst(emb.data(), n, dim, ms, u.data(), v.data(), wt.data()) == n - 1);
  {
    std::vector<int32_t> bu(n), bv(n);
    std::vector<float> bw(n);
    REQUIRE(orc_mst_boruvka(emb.data(), n, dim, ms, bu.data(), bv.data(), bw.data()) == n - 1);
    std::vector<double> e64(emb.begin(), emb.end()), c64(n), w64(n);
    REQUIRE(orc_core_sq_f64(e64.data(), n, dim, ms, c64.data()) == 0);
    REQUIRE(orc_mst_f64(e64.data(), n, dim, ms, bu.data(), bv.data(), w64.data()) == n - 1);
  }
  std::vector<int64_t> left(n - 1), right(n - 1), size(n - 1);
  std::vector<float> ld(n - 1);
  REQUIRE(orc_linkage(u.data(), v.data(), wt.data(), n - 1, n, left.data(), right.data(),
                      ld.data(), size.data()) == 0);
  std::vector<int64_t> par(2 * n), chi(2 * n), csz(2 * n);
  std::vector<float> lam(2 * n);
  const int64_t nn = orc_condense(left.data(), right.data(), ld.data(), size.data(), n - 1, n, 5,
                                  par.data(), chi.data(), lam.data(), csz.data());
  REQUIRE(nn > 0);
  std::vector<int64_t> lv(n);
  const int64_t nl = orc_extract_leaves(par.data(), chi.data(), csz.data(), nn, n, lv.data());
  REQUIRE(nl > 0);
  std::vector<int64_t> labels(n);
  std::vector<float> st(n);
  REQUIRE(orc_label_vector(par.data(), chi.data(), nn, lv.data(), nl, n, labels.data()) == 0);
  REQUIRE(orc_strengths(par.data(), chi.data(), lam.data(), csz.data(), nn, lv.data(), nl,
                        labels.data(), n, st.data()) == 0);
  for (int fp64 = 0; fp64 < 2; ++fp64) {
    std::vector<double> sizes(2 * n), pers(2 * n), births(2 * n), deaths(2 * n), ldeath(2 * n);
    const int64_t r = orc_persistence(par.data(), chi.data(), lam.data(), csz.data(), nn, n, 5,
                                      fp64, sizes.data(), pers.data(), births.data(),
                                      deaths.data(), ldeath.data());
    const int64_t nbc = r >> 32, nbins = r & 0xFFFFFFFF;
    REQUIRE(nbc > 0 && nbins > 0 && nbins <= 2 * n && nbc <= 2 * n);
    std::vector<int64_t> peaks(2 * n + 1), selected(16);
    const int64_t npk = orc_find_peaks(pers.data(), nbins, peaks.data());
    int64_t nsel = 0;
    REQUIRE(orc_select_diverse_peaks(peaks.data(), npk, pers.data(), sizes.data(), births.data(),
                                     deaths.data(), nbc, 0.2, 10, selected.data(), &nsel) == 0);
  }

  // layers and the n-cluster search, f32 (both persistence modes) and f64, both MSTs
  STEP("layers");
  for (int algo = 0; algo < 2; ++algo) {
    orc_set_mst_algo(algo);
    std::vector<int64_t> ll(10 * n);
    std::vector<float> ls(10 * n);
    std::vector<double> lp(10);
    for (int fp64 = 0; fp64 < 2; ++fp64)
      REQUIRE(orc_cluster_layers(emb.data(), n, dim, ms, 5, 0.2, 10, fp64, ll.data(), ls.data(),
                                 lp.data()) > 0);
    std::vector<double> e64(emb.begin(), emb.end()), ls64(10 * n);
    REQUIRE(orc_cluster_layers_f64(e64.data(), n, dim, ms, 5, 0.2, 10, ll.data(), ls64.data(),
                                   lp.data()) > 0);
    REQUIRE(orc_search_n_clusters(emb.data(), n, dim, ms, 8, ll.data(), ls.data()) == 0);
    REQUIRE(orc_search_n_clusters_f64(e64.data(), n, dim, ms, 8, ll.data(), ls64.data()) == 0);
  }
  orc_set_mst_algo(0);

  // the whole pipeline, with its own kNN and with a precomputed one
  STEP("pipeline");
  orc_params_t p = {15, 0.5f, 50, dim, 1.0f, 1, 5, 5, 0, 0.2, 10, 0, 1};
  std::vector<int64_t> el(10 * n);
  std::vector<float> es(10 * n), eemb(n * dim), einit(n * dim);
  std::vector<double> epers(10), tm(5);
  std::vector<int32_t> eidx(n * k);
  std::vector<float> edist(n * k);
  REQUIRE(orc_evoc(x.data(), n, d, nullptr, nullptr, &p, 42, el.data(), es.data(), epers.data(),
                   eidx.data(), edist.data(), eemb.data(), einit.data(), tm.data()) > 0);
  REQUIRE(orc_evoc(x.data(), n, d, eidx.data(), edist.data(), &p, 42, el.data(), es.data(),
                   epers.data(), nullptr, nullptr, nullptr, nullptr, nullptr) > 0);
  p.approx_n_clusters = 6;
  REQUIRE(orc_evoc(x.data(), n, d, nullptr, nullptr, &p, 42, el.data(), es.data(), epers.data(),
                   nullptr, nullptr, nullptr, nullptr, nullptr) > 0);
  std::printf("oracle stages ran clean under the sanitizers\n");
  return 0;
}
