// The condensed cluster tree on the device (src/clustering/condensed_tree.rs:127-248), and
// the per-point labels and membership strengths of several layers at once
// (src/clustering/condensed_tree.rs:311-484, src/clustering/persistence.rs:339-463).
//
// The reference condenses by one depth-first walk from the root: right child first,
// a split into two children of at least min_cluster_size points opens two clusters (left
// first), any smaller child's points fall out at the parent's lambda, and the one large
// child of an unbalanced node keeps the parent's label. Every quantity that walk produces
// is a function of a node's ancestors, so here it is computed for all nodes at once:
//
//  - the walk's visiting order: node y's position in the right-first pre-order is the
//    sum, over y and its ancestors below the root, of 1 (+ the right sibling's subtree,
//    2*size - 1 nodes, for a left child); the number of points before y in that order
//    is the same sum of the right siblings' sizes. Both come from pointer jumping
//    (Wyllie) over the parent links, 2^r ancestors per round, independent of
//    min_cluster_size.
//  - which nodes the walk processes: an internal node is reached unless an ancestor was
//    eliminated, i.e. exactly when it holds at least min_cluster_size points (sizes grow
//    towards the root), or is the root.
//  - cluster numbering: splits are numbered in walk order (a prefix sum over pre-order
//    positions); split s opens clusters n + 1 + 2s (left) and n + 2 + 2s (right).
//  - a processed node's label is that of its nearest ancestor-or-self which is the root
//    or a child of a split; a point falls out at the parent P of its highest ancestor
//    with fewer than min_cluster_size points, with P's label and lambda. Both are
//    "nearest marked ancestor" queries, answered by one more pointer-jumping pass.
//  - the walk's output position of every entry (for the full list of rows): a prefix sum
//    of per-node entry counts in pre-order, plus the point's rank in its eliminated
//    subtree (left block before right block when both children are small).
#include "kernels.h"

namespace evoc {
namespace {

constexpr int TT = 256;

struct alignas(16) Jump {
  int32_t nxt, pre, lf, pad;
};

// pointer-jumping state of the marked-ancestor search (a plain struct: HIP's vector types
// are unions, which the optimiser has mishandled in a conditional initialisation here)
struct alignas(8) Mark {
  int32_t nxt, mark;
};

inline unsigned grid_of(int64_t m) { return (unsigned)std::max<int64_t>(ceil_div(m, TT), 1); }

__device__ inline int32_t node_size(const int32_t* __restrict__ size, int64_t n, int32_t y) {
  return y < n ? 1 : size[y - n];
}

__device__ inline float lambda_of(float d) { return d > 0.0f ? 1.0f / d : INFINITY; }

__global__ __launch_bounds__(TT) void tree_parent_kernel(const int32_t* __restrict__ left,
                                                         const int32_t* __restrict__ right,
                                                         int64_t n, int64_t nr,
                                                         int32_t* __restrict__ par) {
  const int64_t i = (int64_t)blockIdx.x * TT + threadIdx.x;
  if (i >= nr) return;
  par[left[i]] = (int32_t)(n + i);
  par[right[i]] = (int32_t)(n + i);
  if (i == 0) par[2 * nr] = -1;  // the root is nobody's child
}

// increments of the pre-order position and of the points-before count
__global__ __launch_bounds__(TT) void tree_jump_init(const int32_t* __restrict__ par,
                                                     const int32_t* __restrict__ left,
                                                     const int32_t* __restrict__ right,
                                                     const int32_t* __restrict__ size,
                                                     int64_t n, int64_t N,
                                                     Jump* __restrict__ st) {
  const int64_t y = (int64_t)blockIdx.x * TT + threadIdx.x;
  if (y >= N) return;
  const int32_t p = par[y];
  Jump j{p, 0, 0, 0};
  if (p >= 0) {
    if (left[p - n] == (int32_t)y) {
      const int32_t ss = node_size(size, n, right[p - n]);
      j.pre = 2 * ss;  // 1 + (2 * ss - 1)
      j.lf = ss;
    } else {
      j.pre = 1;
    }
  }
  st[y] = j;
}

__global__ __launch_bounds__(TT) void tree_jump_round(const Jump* __restrict__ in, int64_t N,
                                                      Jump* __restrict__ out) {
  const int64_t y = (int64_t)blockIdx.x * TT + threadIdx.x;
  if (y >= N) return;
  Jump v = in[y];
  if (v.nxt >= 0) {
    const Jump w = in[v.nxt];
    v.pre += w.pre;
    v.lf += w.lf;
    v.nxt = w.nxt;
  }
  out[y] = v;
}

__global__ __launch_bounds__(TT) void tree_jump_unpack(const Jump* __restrict__ st, int64_t N,
                                                       int32_t* __restrict__ pre,
                                                       int32_t* __restrict__ lf) {
  const int64_t y = (int64_t)blockIdx.x * TT + threadIdx.x;
  if (y >= N) return;
  const Jump v = st[y];
  pre[y] = v.pre;
  lf[y] = v.lf;
}

// per node at its pre-order position: (is split) << 32 | entries the walk emits there
__global__ __launch_bounds__(TT) void tree_count_kernel(
    const int32_t* __restrict__ left, const int32_t* __restrict__ right,
    const int32_t* __restrict__ size, const int32_t* __restrict__ pre, int64_t n, int64_t N,
    int64_t mcs, int64_t* __restrict__ at) {
  const int64_t y = (int64_t)blockIdx.x * TT + threadIdx.x;
  if (y >= N) return;
  int64_t v = 0;
  if (y >= n) {
    const int64_t r = y - n;
    const bool live = size[r] >= mcs || y == N - 1;
    if (live) {
      const int64_t ls = node_size(size, n, left[r]), rs = node_size(size, n, right[r]);
      const bool lb = ls >= mcs, rb = rs >= mcs;
      if (lb && rb) v = (int64_t(1) << 32) | 2;
      else v = (lb ? 0 : ls) + (rb ? 0 : rs);
    }
  }
  at[pre[y]] = v;
}

// nearest marked ancestor-or-self: marked are the root, the children of splits (label
// sources of processed nodes) and the highest small node on every path (where a point
// falls out); no node is both
__global__ __launch_bounds__(TT) void tree_mark_init(const int32_t* __restrict__ par,
                                                     const int32_t* __restrict__ left,
                                                     const int32_t* __restrict__ right,
                                                     const int32_t* __restrict__ size,
                                                     int64_t n, int64_t N, int64_t mcs,
                                                     Mark* __restrict__ st) {
  const int64_t y = (int64_t)blockIdx.x * TT + threadIdx.x;
  if (y >= N) return;
  const int32_t p = par[y];
  bool mark;
  if (p < 0) {
    mark = true;
  } else {
    const int64_t r = p - n;
    const bool split = node_size(size, n, left[r]) >= mcs && node_size(size, n, right[r]) >= mcs;
    const bool plive = size[r] >= mcs || p == N - 1;
    const bool small = node_size(size, n, (int32_t)y) < mcs;
    mark = split || (small && plive);
  }
  Mark m;
  m.nxt = mark ? -1 : p;
  m.mark = mark ? (int32_t)y : -1;
  st[y] = m;
}

__global__ __launch_bounds__(TT) void tree_mark_round(const Mark* __restrict__ in, int64_t N,
                                                      Mark* __restrict__ out) {
  const int64_t y = (int64_t)blockIdx.x * TT + threadIdx.x;
  if (y >= N) return;
  Mark v = in[y];
  if (v.mark < 0 && v.nxt >= 0) v = in[v.nxt];
  out[y] = v;
}

struct TreeView {
  const int32_t *left, *right, *size, *par, *pre, *lf;
  const float* dist;
  const int64_t* off;  // exclusive prefix of tree_count_kernel's values, by pre-order position
  const Mark* mk;      // .mark: nearest marked ancestor-or-self
  int64_t n, N, mcs;
};

__device__ inline int32_t split_rank(const TreeView& t, int32_t x) {
  return (int32_t)(t.off[t.pre[x]] >> 32);
}

// the cluster label of marked node z (root or a child of a split)
__device__ inline int64_t label_of(const TreeView& t, int32_t z) {
  if (z == (int32_t)(t.N - 1)) return t.n;
  const int32_t p = t.par[z];
  return t.n + 1 + 2 * (int64_t)split_rank(t, p) + (t.left[p - t.n] == z ? 0 : 1);
}

// one thread per node: the split records (and entries) of split nodes, the parent cluster
// and lambda of every point (and its entry)
__global__ __launch_bounds__(TT) void tree_emit_kernel(
    TreeView t, int32_t* __restrict__ pcl, float* __restrict__ plam,
    int32_t* __restrict__ s_par, float* __restrict__ s_lam, int32_t* __restrict__ s_ls,
    int32_t* __restrict__ s_rs, int64_t* __restrict__ e_par, int64_t* __restrict__ e_child,
    float* __restrict__ e_lam, int64_t* __restrict__ e_size) {
  const int64_t y = (int64_t)blockIdx.x * TT + threadIdx.x;
  if (y >= t.N) return;
  const int64_t n = t.n, mcs = t.mcs;
  if (y >= n) {
    const int64_t r = y - n;
    const int32_t ls = node_size(t.size, n, t.left[r]), rs = node_size(t.size, n, t.right[r]);
    if (!(ls >= mcs && rs >= mcs)) return;
    const int64_t o = t.off[t.pre[y]];
    const int32_t sr = (int32_t)(o >> 32);
    const int64_t at = o & 0xffffffffll;
    const int64_t lab = label_of(t, t.mk[y].mark);
    const float lam = lambda_of(t.dist[r]);
    s_par[sr] = (int32_t)lab;
    s_lam[sr] = lam;
    s_ls[sr] = ls;
    s_rs[sr] = rs;
    if (e_par) {
      e_par[at] = lab;
      e_child[at] = n + 1 + 2 * (int64_t)sr;
      e_lam[at] = lam;
      e_size[at] = ls;
      e_par[at + 1] = lab;
      e_child[at + 1] = n + 2 + 2 * (int64_t)sr;
      e_lam[at + 1] = lam;
      e_size[at + 1] = rs;
    }
    return;
  }
  if (mcs <= 1) return;  // every point is a cluster of its own: no point falls out
  const int32_t a = t.mk[y].mark;  // highest small ancestor-or-self
  const int32_t P = t.par[a];
  const int64_t rP = P - n;
  const int64_t lab = label_of(t, t.mk[P].mark);
  const float lam = lambda_of(t.dist[rP]);
  pcl[y] = (int32_t)lab;
  plam[y] = lam;
  if (e_par) {
    const int32_t ls = node_size(t.size, n, t.left[rP]), rs = node_size(t.size, n, t.right[rP]);
    const bool both = ls < mcs && rs < mcs;
    const bool is_right = t.left[rP] != a;
    const int64_t at = (t.off[t.pre[P]] & 0xffffffffll) + ((both && is_right) ? ls : 0) +
                       (t.lf[y] - t.lf[a]);
    e_par[at] = lab;
    e_child[at] = y;
    e_lam[at] = lam;
    e_size[at] = 1;
  }
}

__global__ void tree_total_kernel(const int64_t* __restrict__ at, const int64_t* __restrict__ off,
                                  int64_t N, int64_t* __restrict__ tot) {
  tot[0] = off[N - 1] + at[N - 1];
}

__device__ inline int wave_max(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ void tree_layer_init(int nl, int64_t nd, unsigned* __restrict__ death,
                                int* __restrict__ stats) {
  const int64_t i = (int64_t)blockIdx.x * TT + threadIdx.x;
  if (i < nd) death[i] = 0u;
  if (i < nl) {
    stats[2 * i] = -1;     // largest label
    stats[2 * i + 1] = 0;  // points with a label
  }
}

// the largest lambda at which a point leaves each selected cluster directly
__global__ __launch_bounds__(TT) void tree_death_kernel(const int32_t* __restrict__ pcl,
                                                        const float* __restrict__ plam,
                                                        int64_t n, int nl, int64_t C,
                                                        const int32_t* __restrict__ dsel,
                                                        const int64_t* __restrict__ doff,
                                                        unsigned* __restrict__ death) {
  const int64_t p = (int64_t)blockIdx.x * TT + threadIdx.x;
  if (p >= n) return;
  const int64_t c = pcl[p] - n;
  const unsigned lb = __float_as_uint(plam[p]);  // lambdas are >= 0: bit order is value order
  for (int l = 0; l < nl; ++l) {
    const int32_t r = dsel[l * C + c];
    if (r >= 0) atomicMax(&death[doff[l] + r], lb);
  }
}

__global__ __launch_bounds__(TT) void tree_labels_kernel(
    const int32_t* __restrict__ pcl, const float* __restrict__ plam, int64_t n, int nl,
    int64_t C, const int32_t* __restrict__ clab, const int64_t* __restrict__ doff,
    const unsigned* __restrict__ death, int64_t* __restrict__ labels,
    float* __restrict__ strengths, int* __restrict__ stats) {
  const int64_t p = (int64_t)blockIdx.x * TT + threadIdx.x;
  const bool have = p < n;
  const int64_t c = have ? pcl[p] - n : 0;
  const float lam = have ? plam[p] : 0.0f;
  for (int l = 0; l < nl; ++l) {
    const int32_t lb = have ? clab[l * C + c] : -1;
    if (have) {
      float st = 0.0f;
      if (lb >= 0) {
        const float ml = __uint_as_float(death[doff[l] + lb]);
        st = (ml == 0.0f || !__builtin_isfinite(lam)) ? 1.0f : (lam < ml ? lam : ml) / ml;
      }
      labels[l * n + p] = lb;
      strengths[l * n + p] = st;
    }
    const int mx = wave_max(lb);
    const int cnt = wave_sum(lb >= 0 ? 1 : 0);
    if ((threadIdx.x & 63) == 0) {
      atomicMax(&stats[2 * l], mx);
      if (cnt) atomicAdd(&stats[2 * l + 1], cnt);
    }
  }
}

}  // namespace

void tree_upload(const int64_t* left, const int64_t* right, const float* dist,
                 const int64_t* size, int64_t n, DevTree& t, hipStream_t s, DevPool& pool) {
  EVOC_CHECK(n >= 1 && n <= (int64_t(1) << 30), "tree: n must be in [1, 2^30]");
  t.n = n;
  t.nr = n - 1;
  const int64_t nr = t.nr, N = 2 * n - 1;
  if (nr == 0) return;
  static thread_local std::vector<int32_t> h;  // kept between calls: no page faults
  h.resize(3 * nr);
  for (int64_t i = 0; i < nr; ++i) {
    h[i] = (int32_t)left[i];
    h[nr + i] = (int32_t)right[i];
    h[2 * nr + i] = (int32_t)size[i];
  }
  t.rows.alloc(pool, 3 * nr);
  t.dist.alloc(pool, nr);
  EVOC_HIP(hipMemcpyAsync(t.rows.p, h.data(), 3 * nr * sizeof(int32_t), hipMemcpyHostToDevice, s));
  EVOC_HIP(hipMemcpyAsync(t.dist.p, dist, nr * sizeof(float), hipMemcpyHostToDevice, s));
  t.par.alloc(pool, N);
  t.pre.alloc(pool, N);
  t.lf.alloc(pool, N);
  tree_parent_kernel<<<grid_of(nr), TT, 0, s>>>(t.left(), t.right(), n, nr, t.par.p);
  Buf<Jump> a(pool, N), b(pool, N);
  tree_jump_init<<<grid_of(N), TT, 0, s>>>(t.par.p, t.left(), t.right(), t.size(), n, N, a.p);
  // 2^rounds > N > the deepest path: every node has reached the root
  int rounds = 0;
  while ((int64_t(1) << rounds) <= N) ++rounds;
  for (int r = 0; r < rounds; ++r) {
    tree_jump_round<<<grid_of(N), TT, 0, s>>>(a.p, N, b.p);
    std::swap(a.p, b.p);
  }
  tree_jump_unpack<<<grid_of(N), TT, 0, s>>>(a.p, N, t.pre.p, t.lf.p);
  EVOC_HIP(hipGetLastError());
  EVOC_HIP(hipStreamSynchronize(s));  // h is a pageable host buffer
}

void tree_condense(const DevTree& t, int64_t mcs, DevCondensed& out, CondensedRows* full,
                   hipStream_t s, DevPool& pool) {
  const int64_t n = t.n, nr = t.nr, N = 2 * n - 1;
  out.n = n;
  out.mcs = mcs;
  out.S = 0;
  out.E = 0;
  out.split_parent.clear();
  out.split_lam.clear();
  out.split_ls.clear();
  out.split_rs.clear();
  if (full) *full = CondensedRows{};
  if (nr == 0) return;
  Buf<int64_t> at(pool, N), off(pool, N), tot(pool, 1);
  tree_count_kernel<<<grid_of(N), TT, 0, s>>>(t.left(), t.right(), t.size(), t.pre.p, n, N,
                                              mcs, at.p);
  exclusive_scan_i64(at.p, off.p, N, s, pool);
  tree_total_kernel<<<1, 1, 0, s>>>(at.p, off.p, N, tot.p);
  Buf<Mark> ma(pool, N), mb(pool, N);
  tree_mark_init<<<grid_of(N), TT, 0, s>>>(t.par.p, t.left(), t.right(), t.size(), n, N, mcs,
                                           ma.p);
  int rounds = 0;
  while ((int64_t(1) << rounds) <= N) ++rounds;
  for (int r = 0; r < rounds; ++r) {
    tree_mark_round<<<grid_of(N), TT, 0, s>>>(ma.p, N, mb.p);
    std::swap(ma.p, mb.p);
  }
  EVOC_HIP(hipGetLastError());
  int64_t total = 0;
  EVOC_HIP(hipMemcpyAsync(&total, tot.p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
  const int64_t S = total >> 32, E = total & 0xffffffffll;
  out.S = S;
  out.E = E;
  out.pcl.alloc(pool, n);
  out.plam.alloc(pool, n);
  Buf<int32_t> sp(pool, std::max<int64_t>(S, 1) * 3);
  Buf<float> sl(pool, std::max<int64_t>(S, 1));
  Buf<int64_t> ep, ec, es;
  Buf<float> el;
  if (full) {
    ep.alloc(pool, E);
    ec.alloc(pool, E);
    es.alloc(pool, E);
    el.alloc(pool, E);
  }
  TreeView v{t.left(), t.right(), t.size(), t.par.p, t.pre.p, t.lf.p, t.dist.p, off.p, ma.p,
             n, N, mcs};
  tree_emit_kernel<<<grid_of(N), TT, 0, s>>>(v, out.pcl.p, out.plam.p, sp.p, sl.p, sp.p + S,
                                             sp.p + 2 * S, full ? ep.p : nullptr,
                                             full ? ec.p : nullptr, full ? el.p : nullptr,
                                             full ? es.p : nullptr);
  EVOC_HIP(hipGetLastError());
  out.split_parent.resize(S);
  out.split_lam.resize(S);
  out.split_ls.resize(S);
  out.split_rs.resize(S);
  if (S > 0) {
    EVOC_HIP(hipMemcpyAsync(out.split_parent.data(), sp.p, S * sizeof(int32_t),
                            hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(out.split_ls.data(), sp.p + S, S * sizeof(int32_t),
                            hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(out.split_rs.data(), sp.p + 2 * S, S * sizeof(int32_t),
                            hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(out.split_lam.data(), sl.p, S * sizeof(float),
                            hipMemcpyDeviceToHost, s));
  }
  if (full) {
    full->parent.resize(E);
    full->child.resize(E);
    full->size.resize(E);
    full->lambda.resize(E);
    EVOC_HIP(hipMemcpyAsync(full->parent.data(), ep.p, E * sizeof(int64_t),
                            hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(full->child.data(), ec.p, E * sizeof(int64_t),
                            hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(full->size.data(), es.p, E * sizeof(int64_t),
                            hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(full->lambda.data(), el.p, E * sizeof(float),
                            hipMemcpyDeviceToHost, s));
  }
  EVOC_HIP(hipStreamSynchronize(s));
}

void tree_layer_labels(const DevCondensed& c, int nl, const std::vector<int32_t>& clab,
                       const std::vector<int32_t>& dsel, const std::vector<int64_t>& nsel,
                       int64_t* labels, float* strengths, std::vector<int64_t>& order,
                       std::vector<int64_t>& ncl, std::vector<int64_t>& assigned, hipStream_t s,
                       DevPool& pool) {
  const int64_t n = c.n, C = 2 * c.S + 1;
  EVOC_CHECK(c.mcs >= 2, "tree_layer_labels: needs min_cluster_size >= 2");
  EVOC_CHECK((int64_t)clab.size() == nl * C && (int64_t)dsel.size() == nl * C &&
                 (int64_t)nsel.size() == nl,
             "tree_layer_labels: table sizes");
  order.clear();
  ncl.assign(nl, 0);
  assigned.assign(nl, 0);
  if (nl == 0 || n == 0) return;
  std::vector<int64_t> doff(nl);
  int64_t nd = 0;
  for (int l = 0; l < nl; ++l) {
    doff[l] = nd;
    nd += nsel[l];
  }
  Buf<int32_t> dclab(pool, nl * C), ddsel(pool, nl * C);
  Buf<int64_t> ddoff(pool, nl), dlab(pool, (size_t)nl * n);
  Buf<unsigned> death(pool, std::max<int64_t>(nd, 1));
  Buf<float> dstr(pool, (size_t)nl * n);
  Buf<int> dstats(pool, 2 * nl);
  EVOC_HIP(hipMemcpyAsync(dclab.p, clab.data(), nl * C * sizeof(int32_t), hipMemcpyHostToDevice, s));
  EVOC_HIP(hipMemcpyAsync(ddsel.p, dsel.data(), nl * C * sizeof(int32_t), hipMemcpyHostToDevice, s));
  EVOC_HIP(hipMemcpyAsync(ddoff.p, doff.data(), nl * sizeof(int64_t), hipMemcpyHostToDevice, s));
  tree_layer_init<<<grid_of(std::max<int64_t>(nd, nl)), TT, 0, s>>>(nl, nd, death.p, dstats.p);
  tree_death_kernel<<<grid_of(n), TT, 0, s>>>(c.pcl.p, c.plam.p, n, nl, C, ddsel.p, ddoff.p,
                                              death.p);
  tree_labels_kernel<<<grid_of(n), TT, 0, s>>>(c.pcl.p, c.plam.p, n, nl, C, dclab.p, ddoff.p,
                                               death.p, dlab.p, dstr.p, dstats.p);
  EVOC_HIP(hipGetLastError());
  std::vector<int> stats(2 * nl);
  EVOC_HIP(hipMemcpyAsync(stats.data(), dstats.p, 2 * nl * sizeof(int), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
  for (int l = 0; l < nl; ++l) {
    ncl[l] = (int64_t)stats[2 * l] + 1;
    assigned[l] = stats[2 * l + 1];
  }
  order = order_by_clusters(ncl);
  // through the pool's pinned staging buffer: the copy runs at link speed, and the
  // destination's first touch is spread over host threads
  const size_t lb = (size_t)n * sizeof(int64_t), sb = (size_t)n * sizeof(float);
  char* stage = static_cast<char*>(pool.pinned((size_t)nl * (lb + sb)));
  for (int q = 0; q < nl; ++q) {
    const int64_t o = order[q];
    EVOC_HIP(hipMemcpyAsync(stage + q * lb, dlab.p + (size_t)o * n, lb, hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(stage + nl * lb + q * sb, dstr.p + (size_t)o * n, sb,
                            hipMemcpyDeviceToHost, s));
  }
  EVOC_HIP(hipStreamSynchronize(s));
  parallel_copy(labels, stage, nl * lb);
  parallel_copy(strengths, stage + nl * lb, nl * sb);
}

}  // namespace evoc
