// The tree stages of evoc() on the device: the condensed tree
// (src/clustering/condensed_tree.rs:127-248) and the per-point labels and membership
// strengths of every layer (src/clustering/condensed_tree.rs:311-484,
// src/clustering/persistence.rs:339-463) run as kernels (tree.hip); the cluster-level
// steps over the S splits (barcodes, persistence, peaks, the selection of diverse layers,
// the nearest selected ancestor of every cluster) run on the host, on the split records
// the condensation returns. The approx_n_clusters search (src/lib.rs:377-447) counts
// leaves per probe without condensing (search_bounds) and condenses twice at the end.
#include <chrono>
#include <cstdio>
#include <cstring>

#include "host_stages.h"

namespace evoc {

namespace {

void upload(const Linkage& L, DevTree& t, hipStream_t s, DevPool& pool) {
  tree_upload(L.left.data(), L.right.data(), L.dist.data(), L.size.data(), L.n, t, s, pool);
}

CTree to_ctree(CondensedRows&& r, int64_t n) {
  CTree t;
  t.n = n;
  t.parent = std::move(r.parent);
  t.child = std::move(r.child);
  t.size = std::move(r.size);
  t.lambda = std::move(r.lambda);
  return t;
}

// one layer of a condensation: its leaves as the selection
struct LeafLayer {
  int64_t nleaves = 0, assigned = 0;
  std::vector<int64_t> labels;
  std::vector<float> strengths;
};

LeafLayer leaf_layer(const DevTree& t, int64_t m, hipStream_t s, DevPool& pool) {
  const int64_t n = t.n;
  LeafLayer out;
  out.labels.assign(n, -1);
  out.strengths.assign(n, 0.0f);
  if (n < 2) return out;  // no linkage rows: empty tree, no leaves
  if (m >= 2) {
    DevCondensed c;
    tree_condense(t, m, c, nullptr, s, pool);
    std::vector<std::vector<int64_t>> sel(1, split_leaves(c));
    out.nleaves = (int64_t)sel[0].size();
    std::vector<int32_t> clab, dsel;
    std::vector<int64_t> nsel, order, ncl, assigned;
    layer_tables(c, sel, clab, dsel, nsel);
    tree_layer_labels(c, 1, clab, dsel, nsel, out.labels.data(), out.strengths.data(), order,
                      ncl, assigned, s, pool);
    out.assigned = assigned[0];
  } else {
    CondensedRows rows;
    DevCondensed c;
    tree_condense(t, m, c, &rows, s, pool);
    const CTree full = to_ctree(std::move(rows), n);
    const std::vector<int64_t> lv = leaves(full);
    out.nleaves = (int64_t)lv.size();
    labels_and_strengths(full, lv, out.labels.data(), out.strengths.data());
    for (auto x : out.labels) out.assigned += x >= 0;
  }
  return out;
}

}  // namespace

CTree condense_device(const Linkage& L, int64_t mcs, hipStream_t s, DevPool& pool) {
  DevTree t;
  upload(L, t, s, pool);
  CondensedRows rows;
  DevCondensed c;
  tree_condense(t, mcs, c, &rows, s, pool);
  return to_ctree(std::move(rows), L.n);
}

Layers cluster_layers_device(const Linkage& L, int64_t base_mcs, double min_sim,
                             int64_t max_layers, hipStream_t s, DevPool& pool, int verbose) {
  const int64_t n = L.n;
  auto tl = std::chrono::steady_clock::now();
  auto lap = [&](const char* what) {
    if (verbose < 2) return;
    const auto t = std::chrono::steady_clock::now();
    std::printf("      %s: %.1f ms\n", what,
                std::chrono::duration<double, std::milli>(t - tl).count());
    tl = t;
  };
  if (n < 2) {  // an empty condensed tree: one layer, nothing labelled
    Layers out(n, 1);
    std::fill(out.labels, out.labels + n, -1);
    std::fill(out.strengths, out.strengths + n, 0.0f);
    return out;
  }
  DevTree t;
  upload(L, t, s, pool);
  lap("tree upload and walk order");
  if (base_mcs < 2) {
    CondensedRows rows;
    DevCondensed c;
    tree_condense(t, base_mcs, c, &rows, s, pool);
    const CTree full = to_ctree(std::move(rows), n);
    return layers_from_rows(full, select_layers(full, leaves(full), base_mcs, min_sim,
                                                max_layers));
  }
  DevCondensed c;
  tree_condense(t, base_mcs, c, nullptr, s, pool);
  lap("condense");
  const LayerSelection sel =
      select_layers(split_rows(c), split_leaves(c), base_mcs, min_sim, max_layers);
  lap("layer selection");
  const int nl = (int)sel.sel.size();
  std::vector<int32_t> clab, dsel;
  std::vector<int64_t> nsel, order, ncl, assigned;
  layer_tables(c, sel.sel, clab, dsel, nsel);
  Layers out(n, nl);
  tree_layer_labels(c, nl, clab, dsel, nsel, out.labels, out.strengths, order, ncl, assigned, s,
                    pool);
  for (int q = 0; q < nl; ++q) out.persistence[q] = sel.pers[order[q]];
  lap("labels and strengths");
  return out;
}

void search_n_clusters_device(const Linkage& L, int64_t target_k, int64_t* labels,
                              float* strengths, hipStream_t s, DevPool& pool) {
  const int64_t n = L.n;
  int64_t lo = 2, hi = n / 2;
  search_bounds(L, target_k, lo, hi);
  DevTree t;
  if (n >= 2) upload(L, t, s, pool);
  else t.n = n;
  const LeafLayer a = leaf_layer(t, lo, s, pool);
  const LeafLayer b = leaf_layer(t, hi, s, pool);
  const int64_t dl = std::llabs(a.nleaves - target_k), dh = std::llabs(b.nleaves - target_k);
  const bool use_lo = dl < dh || (dl == dh && a.assigned >= b.assigned);
  const LeafLayer& r = use_lo ? a : b;
  std::memcpy(labels, r.labels.data(), n * sizeof(int64_t));
  std::memcpy(strengths, r.strengths.data(), n * sizeof(float));
}

}  // namespace evoc
