// Core distances and the mutual-reachability MST (build_mst, src/clustering/mst.rs:82-166)
// on gfx950, with a kd-tree of its own layout in place of src/clustering/kd_tree.rs: the
// embedding is cut into 64-point tiles, the tiles are the leaves of an implicit binary
// tree in heap order, and level by level every node's points are ordered along the
// widest dimension of its box, so that the count split at its middle tile is a median
// split. Every query wavefront (one tile of 64 queries) walks that tree depth first,
// nearer child first, over a wave-uniform stack, skipping a subtree when every lane's
// point-to-box distance already exceeds its bound (the AABB rule of kd_tree.rs:330-339)
// or when the subtree lies wholly in the lane's component (same-component pruning,
// kd_tree.rs:336-339).
//
// Ties: core distances are order-free; MST edges are chosen under the strict total
// order (w^2, min(u,v), max(u,v)) on original point ids, so the tree is the unique
// minimum spanning tree for that order. Per Boruvka round each component's minimum
// outgoing edge is reduced with two integer atomicMin passes (weight bits, then ids).
#include <cstring>
#include <numeric>

#include "kernels.h"

namespace evoc {
namespace {

constexpr int TS = 64;  // points per tile (one wavefront of queries)
constexpr int MAXD = 16;
constexpr int KCORE = 32;     // core-distance list of the common kernel
constexpr int KCORE_MAX = 64;  // min_samples above KCORE take a 64-entry list

// kd ordering: positions are cut into 64-point tiles, and tile t is leaf L + t of the
// implicit tree. Level by level, every node's points are ordered along the widest
// dimension of the node's box, so that the fixed count split at the node's middle tile
// boundary is a median split (a balanced kd-tree in the tree's own heap layout).
__global__ void perm_tile_box_kernel(const float* __restrict__ e, const uint32_t* __restrict__ perm,
                                     int64_t n, int dim, int64_t ntiles, float* __restrict__ blo,
                                     float* __restrict__ bhi) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t t = (int64_t)blockIdx.x * 4 + wave;
  if (t >= ntiles) return;
  const int64_t p = t * TS + lane;
  const int64_t row = p < n ? (int64_t)perm[p] : 0;
  for (int c = 0; c < dim; ++c) {
    float x = p < n ? e[row * dim + c] : 0.0f;
    if (!(x == x)) x = 0.0f;
    float l = p < n ? x : __builtin_huge_valf();
    float h = p < n ? x : -__builtin_huge_valf();
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      l = fminf(l, __shfl_xor(l, o, 64));
      h = fmaxf(h, __shfl_xor(h, o, 64));
    }
    if (lane == 0) {
      blo[t * dim + c] = l;
      bhi[t * dim + c] = h;
    }
  }
}

// one workgroup per node of level `lev`: the box of the tiles under it
__global__ __launch_bounds__(256) void kd_node_box_kernel(const float* __restrict__ blo,
                                                          const float* __restrict__ bhi,
                                                          int64_t ntiles, int64_t span, int dim,
                                                          float* __restrict__ olo,
                                                          float* __restrict__ ohi) {
  __shared__ float rl[4], rh[4];
  const int64_t node = blockIdx.x;
  const int64_t t0 = node * span, t1 = min<int64_t>(t0 + span, ntiles);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c = 0; c < dim; ++c) {
    float l = __builtin_huge_valf(), h = -__builtin_huge_valf();
    for (int64_t t = t0 + threadIdx.x; t < t1; t += 256) {
      l = fminf(l, blo[t * dim + c]);
      h = fmaxf(h, bhi[t * dim + c]);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      l = fminf(l, __shfl_xor(l, o, 64));
      h = fmaxf(h, __shfl_xor(h, o, 64));
    }
    if (lane == 0) {
      rl[wave] = l;
      rh[wave] = h;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      olo[node * dim + c] = fminf(fminf(rl[0], rl[1]), fminf(rl[2], rl[3]));
      ohi[node * dim + c] = fmaxf(fmaxf(rh[0], rh[1]), fmaxf(rh[2], rh[3]));
    }
    __syncthreads();
  }
}

// sort key of position p at level `lev`: (node, coordinate along the node's widest
// dimension quantised to 16 bits over the node's extent); the stable sort keeps the
// previous order among equal keys
__global__ void kd_key_kernel(const float* __restrict__ e, const uint32_t* __restrict__ perm,
                              int64_t n, int dim, int shift, const float* __restrict__ olo,
                              const float* __restrict__ ohi, uint64_t* __restrict__ key,
                              uint32_t* __restrict__ val) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int64_t node = (p / TS) >> shift;
  int best = 0;
  float wb = ohi[node * dim] - olo[node * dim];
  for (int c = 1; c < dim; ++c) {
    const float w = ohi[node * dim + c] - olo[node * dim + c];
    if (w > wb) {
      wb = w;
      best = c;
    }
  }
  const uint32_t row = perm[p];
  float x = e[(int64_t)row * dim + best];
  if (!(x == x)) x = 0.0f;
  const float lo = olo[node * dim + best];
  const float r = wb > 0.0f ? (x - lo) / wb : 0.0f;
  const uint32_t q = (uint32_t)fminf(fmaxf(r * 65535.0f, 0.0f), 65535.0f);
  key[p] = ((uint64_t)node << 16) | q;
  val[p] = row;
}

__global__ void iota_kernel(int64_t n, uint32_t* __restrict__ v) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) v[p] = (uint32_t)p;
}

__global__ void gather_kernel(const float* __restrict__ e, const uint32_t* __restrict__ perm,
                              int64_t n, int dim, float* __restrict__ es) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * dim) return;
  const int64_t p = q / dim;
  const int c = (int)(q - p * dim);
  es[q] = e[(int64_t)perm[p] * dim + c];
}

// one wavefront per tile: bounding box (lo, hi) per dimension
__global__ void tile_box_kernel(const float* __restrict__ es, int64_t n, int dim, int64_t ntiles,
                                float* __restrict__ blo, float* __restrict__ bhi) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t t = (int64_t)blockIdx.x * 4 + wave;
  if (t >= ntiles) return;
  const int64_t p = t * TS + lane;
  for (int c = 0; c < dim; ++c) {
    float l = p < n ? es[p * dim + c] : __builtin_huge_valf();
    float h = p < n ? es[p * dim + c] : -__builtin_huge_valf();
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      l = fminf(l, __shfl_xor(l, o, 64));
      h = fmaxf(h, __shfl_xor(h, o, 64));
    }
    if (lane == 0) {
      blo[t * dim + c] = l;
      bhi[t * dim + c] = h;
    }
  }
}

template <int D>
__device__ __forceinline__ float box_dist2(const float (&q)[MAXD], const float* __restrict__ lo,
                                           const float* __restrict__ hi, int dim) {
  float r = 0.0f;
#pragma unroll
  for (int c = 0; c < MAXD; ++c)
    if (c < (D > 0 ? D : dim)) {
      const float l = lo[c], h = hi[c], x = q[c];
      const float df = x < l ? l - x : (x > h ? x - h : 0.0f);
      r = r + df * df;
    }
  return r;
}

template <int D>
__device__ __forceinline__ float dsq_seq(const float (&q)[MAXD], const float* __restrict__ b,
                                         int dim) {
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < MAXD; ++c)
    if (c < (D > 0 ? D : dim)) {
      const float df = q[c] - b[c];
      const float p = df * df;
      s = s + p;
    }
  return s;
}

constexpr int STACK = 64;  // > tree depth + 1 for any n < 2^31
// a tile needed by more queries of a wavefront than this is scanned lane per query
constexpr int CORE_DIRECT = 24;
constexpr int SEARCH_DIRECT = 24;

__device__ __forceinline__ float wave_min(float x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x = fminf(x, __shfl_xor(x, o, 64));
  return x;
}

// Lowers each lane's component bound to its w (a real edge weight; +inf: nothing to
// publish). Called by the whole wavefront. In late rounds a few large components take
// every wavefront's updates, and one atomic per lane serialises on their addresses, so
// the wavefront sends one atomic per distinct component among its lanes (usually one:
// tiles are compact), carrying that component's minimum.
__device__ __forceinline__ void publish_bound(unsigned* __restrict__ comp_w, bool valid,
                                              int32_t mycomp, float w) {
  const int lane = threadIdx.x % 64;
  bool pend = valid && w < __builtin_huge_valf();
  for (unsigned long long m = __ballot(pend); m != 0; m = __ballot(pend)) {
    const int lead = __ffsll((long long)m) - 1;
    const int32_t c = __shfl(mycomp, lead, 64);
    const float cm = wave_min((pend && mycomp == c) ? w : __builtin_huge_valf());
    if (lane == lead) atomicMin(&comp_w[c], __float_as_uint(cm));
    pend = pend && mycomp != c;
  }
}

// leaves L..2L-1 hold the tile boxes (empty boxes past the last tile)
__global__ void bvh_leaf_kernel(const float* __restrict__ tlo, const float* __restrict__ thi,
                                int64_t ntiles, int64_t L, int dim, float* __restrict__ nlo,
                                float* __restrict__ nhi) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L) return;
  for (int c = 0; c < dim; ++c) {
    nlo[(L + t) * dim + c] = t < ntiles ? tlo[t * dim + c] : __builtin_huge_valf();
    nhi[(L + t) * dim + c] = t < ntiles ? thi[t * dim + c] : -__builtin_huge_valf();
  }
}
__global__ void bvh_merge_kernel(int64_t lev, int dim, float* __restrict__ nlo,
                                 float* __restrict__ nhi) {
  const int64_t v = lev + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= 2 * lev) return;
  for (int c = 0; c < dim; ++c) {
    nlo[v * dim + c] = fminf(nlo[2 * v * dim + c], nlo[(2 * v + 1) * dim + c]);
    nhi[v * dim + c] = fmaxf(nhi[2 * v * dim + c], nhi[(2 * v + 1) * dim + c]);
  }
}
// component of a subtree: c if all its points are in component c, -1 if mixed, -2 empty
__global__ void bvh_comp_leaf_kernel(const int32_t* __restrict__ tile_comp, int64_t ntiles,
                                     int64_t L, int32_t* __restrict__ ncomp) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < L) ncomp[L + t] = t < ntiles ? tile_comp[t] : -2;
}
__global__ void bvh_comp_merge_kernel(int64_t lev, int32_t* __restrict__ ncomp) {
  const int64_t v = lev + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= 2 * lev) return;
  const int32_t a = ncomp[2 * v], b = ncomp[2 * v + 1];
  ncomp[v] = a == -2 ? b : (b == -2 ? a : (a == b ? a : -1));
}

__device__ __forceinline__ float rdlane(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const unsigned lo = __shfl_xor((unsigned)x, o, 64), hi = __shfl_xor((unsigned)(x >> 32), o, 64);
    const unsigned long long y = ((unsigned long long)hi << 32) | lo;
    x = y < x ? y : x;
  }
  return x;
}

// ------------------------------------------------------------------ core distances
// With LIST, the kernel also records each point's k nearest (sorted positions) and
// whether another point lies at exactly its k-th distance (a tie outside the list): the
// first Boruvka round resolves most points from these lists.
template <int D, int KC, bool LIST>
__global__ __launch_bounds__(64) void core_kernel(const float* __restrict__ es, int64_t n,
                                                  int dim, int64_t ntiles, int64_t L,
                                                  const float* __restrict__ nlo,
                                                  const float* __restrict__ nhi, int k,
                                                  float* __restrict__ core,
                                                  int32_t* __restrict__ knn_pos,
                                                  uint8_t* __restrict__ knn_tie) {
  const int lane = threadIdx.x;
  const int64_t qt = blockIdx.x;
  const int64_t p = qt * TS + lane;
  const bool valid = p < n;
  const int nd = D > 0 ? D : dim;
  float q[MAXD];
#pragma unroll
  for (int c = 0; c < MAXD; ++c) q[c] = (valid && c < nd) ? es[p * nd + c] : 0.0f;
  // ascending list of the k smallest squared distances; kth = best[k-1]
  float best[KC];
  int32_t bidx[LIST ? KC : 1];
#pragma unroll
  for (int j = 0; j < KC; ++j) best[j] = __builtin_huge_valf();
  if (LIST) {
#pragma unroll
    for (int j = 0; j < KC; ++j) bidx[LIST ? j : 0] = -1;
  }
  float kth = __builtin_huge_valf();
  float rej = __builtin_huge_valf();  // smallest distance not in the list
  // One tile. When many queries need it (the own tile, or early in the walk), each lane
  // scans the tile's points from LDS. Otherwise the scan is transposed: lane j holds the
  // tile's point j, and the queries that need the tile (mine) are taken one at a time; the
  // walk is shared by the wavefront, so most tiles it reaches are needed by a few of its
  // queries only. Either way a query inserts its candidates below its k-th distance in
  // tile order, as a scan of its own would (candidates it skips are farther than its k-th
  // distance, so they could only lower rej by values above it).
  __shared__ float tp[TS * MAXD];
  auto scan_tile = [&](int64_t t, bool mine) {
    const int64_t pt = t * TS + lane;
    const bool cv = pt < n;
    const unsigned long long need0 = __ballot(valid && mine);
    auto insert = [&](float xv, int32_t xi) {
      if (xv < kth) {
        if (LIST) {
#pragma unroll
          for (int m = 0; m < KC; ++m)
            if (m < k) {
              const bool sw = xv < best[m];
              const float lo = sw ? xv : best[m], hi = sw ? best[m] : xv;
              const int32_t li = sw ? xi : bidx[LIST ? m : 0], hi_i = sw ? bidx[LIST ? m : 0] : xi;
              best[m] = lo;
              bidx[LIST ? m : 0] = li;
              xv = hi;
              xi = hi_i;
            }
          rej = fminf(rej, xv);  // the entry pushed out of the list
        } else {
#pragma unroll
          for (int m = 0; m < KC; ++m)
            if (m < k) {
              const float lo = fminf(xv, best[m]);
              xv = fmaxf(xv, best[m]);
              best[m] = lo;
            }
        }
#pragma unroll
        for (int m = 0; m < KC; ++m)
          if (m == k - 1) kth = best[m];
      } else if (LIST) {
        rej = fminf(rej, xv);
      }
    };
    if (__popcll(need0) > CORE_DIRECT) {
      __syncthreads();
      for (int c = 0; c < nd; ++c) tp[lane * nd + c] = cv ? es[pt * nd + c] : 0.0f;
      __syncthreads();
      const int cnt = (int)min<int64_t>(TS, n - t * TS);
      if (valid) {
        for (int j = 0; j < cnt; ++j) {
          if (t * TS + j == p) continue;
          insert(dsq_seq<D>(q, tp + j * nd, nd), (int32_t)(t * TS + j));
        }
      }
      return;
    }
    float cx[MAXD];
#pragma unroll
    for (int c = 0; c < MAXD; ++c) cx[c] = (cv && c < nd) ? es[pt * nd + c] : 0.0f;
    for (unsigned long long need = need0; need != 0; need &= need - 1) {
      const int ql = __ffsll((long long)need) - 1;
      float qv[MAXD];
#pragma unroll
      for (int c = 0; c < MAXD; ++c) qv[c] = c < nd ? rdlane(q[c], ql) : 0.0f;
      const float kq = rdlane(kth, ql);
      const bool el = cv && pt != qt * TS + ql;
      const float x = el ? dsq_seq<D>(qv, cx, nd) : __builtin_huge_valf();
      const bool pass = el && x < kq;
      if (LIST) {
        const float r = wave_min((el && !pass) ? x : __builtin_huge_valf());
        if (lane == ql) rej = fminf(rej, r);
      }
      for (unsigned long long pm = __ballot(pass); pm != 0; pm &= pm - 1) {
        const int jl = __ffsll((long long)pm) - 1;
        const float xv = rdlane(x, jl);
        if (lane == ql) insert(xv, (int32_t)(t * TS + jl));
      }
    }
  };
  scan_tile(qt, true);  // the own tile first, for a starting bound
  // depth-first walk, nearer child first, over a wave-uniform stack in LDS
  __shared__ int32_t stk[STACK];
  int sp = 0;
  stk[sp++] = 1;
  while (sp > 0) {
    const int64_t v = __builtin_amdgcn_readfirstlane(stk[--sp]);  // wave-uniform: scalar loads of the boxes
    const bool far = !valid || box_dist2<D>(q, nlo + v * nd, nhi + v * nd, nd) > kth;
    if (__all(far)) continue;
    if (v >= L) {
      if (v - L != qt) scan_tile(v - L, !far);
      continue;
    }
    const float d0 = box_dist2<D>(q, nlo + 2 * v * nd, nhi + 2 * v * nd, nd);
    const float d1 = box_dist2<D>(q, nlo + (2 * v + 1) * nd, nhi + (2 * v + 1) * nd, nd);
    const bool n0 = !__all(!valid || d0 > kth), n1 = !__all(!valid || d1 > kth);
    const float m0 = wave_min(valid ? d0 : __builtin_huge_valf());
    const float m1 = wave_min(valid ? d1 : __builtin_huge_valf());
    const int32_t near = (int32_t)(m1 < m0 ? 2 * v + 1 : 2 * v);
    const bool nn = m1 < m0 ? n1 : n0, nf = m1 < m0 ? n0 : n1;
    if (nf) stk[sp++] = near ^ 1;
    if (nn) stk[sp++] = near;
  }
  if (valid) {
    core[p] = kth;
    if (LIST) {
#pragma unroll
      for (int m = 0; m < KC; ++m)
        if (m < k) knn_pos[p * k + m] = bidx[LIST ? m : 0];
      knn_tie[p] = rej == kth ? 1 : 0;  // boxes pruned held only points farther than kth
    }
  }
}

// ------------------------------------------------------------------ Boruvka search
struct Best {
  float w;
  int32_t a, b;  // original ids, a < b
};
__device__ __forceinline__ bool better(float w, int32_t a, int32_t b, const Best& o) {
  if (w != o.w) return w < o.w;
  if (a != o.a) return a < o.a;
  return b < o.b;
}

template <int D>
__global__ __launch_bounds__(64) void boruvka_search_kernel(
    const float* __restrict__ es, const float* __restrict__ core, const int32_t* __restrict__ comp,
    const uint32_t* __restrict__ perm, int64_t n, int dim, int64_t L,
    const float* __restrict__ nlo, const float* __restrict__ nhi,
    const int32_t* __restrict__ ncomp, unsigned* __restrict__ comp_w,
    float* __restrict__ best_w, int32_t* __restrict__ best_a, int32_t* __restrict__ best_b,
    const float* __restrict__ pre_w, const int32_t* __restrict__ pre_a,
    const int32_t* __restrict__ pre_b, const uint8_t* __restrict__ pre_done) {
  __shared__ float tp[TS * MAXD];
  __shared__ float tc[TS];
  __shared__ int32_t tcomp[TS];
  __shared__ int32_t tid_[TS];
  const int lane = threadIdx.x;
  const int64_t qt = blockIdx.x;
  const int64_t p = qt * TS + lane;
  const bool valid = p < n;
  const int nd = D > 0 ? D : dim;
  float q[MAXD];
#pragma unroll
  for (int c = 0; c < MAXD; ++c) q[c] = (valid && c < nd) ? es[p * nd + c] : 0.0f;
  const float cq = valid ? core[p] : 0.0f;
  const int32_t mycomp = valid ? comp[p] : -1;
  const int32_t myid = valid ? (int32_t)perm[p] : 0;
  Best bst{__builtin_huge_valf(), INT32_MAX, INT32_MAX};
  // first round: the edge resolved from the core-distance neighbour list (a real edge,
  // so a bound); a point marked done needs no search at all
  bool done = false;
  if (pre_w && valid) {
    bst = Best{pre_w[p], pre_a[p], pre_b[p]};
    done = pre_done[p] != 0;
    if (bst.w < __builtin_huge_valf()) atomicMin(&comp_w[mycomp], __float_as_uint(bst.w));
  }
  const bool live = valid && !done;
  __shared__ int32_t stk[STACK];
  int sp = 0;
  stk[sp++] = 1;
  while (sp > 0) {
    const int64_t v = __builtin_amdgcn_readfirstlane(stk[--sp]);  // wave-uniform: scalar loads of the boxes
    const int32_t vc = ncomp[v];
    const float cw = valid ? __uint_as_float(comp_w[mycomp]) : __builtin_huge_valf();
    float bound = fminf(bst.w, cw);
    const bool prune = !live || vc == -2 || (vc >= 0 && vc == mycomp) || cq > bound ||
                       box_dist2<D>(q, nlo + v * nd, nhi + v * nd, nd) > bound;
    if (__all(prune)) continue;
    if (v >= L) {
      const int64_t t = v - L;
      const int64_t pt = t * TS + lane;
      const bool cv = pt < n;
      const unsigned long long need0 = __ballot(!prune);
      if (__popcll(need0) > SEARCH_DIRECT) {
        __syncthreads();
        for (int c = 0; c < nd; ++c) tp[lane * nd + c] = cv ? es[pt * nd + c] : 0.0f;
        tc[lane] = cv ? core[pt] : __builtin_huge_valf();
        tcomp[lane] = cv ? comp[pt] : mycomp;
        tid_[lane] = cv ? (int32_t)perm[pt] : 0;
        __syncthreads();
        if (!prune) {
          const int cnt = (int)min<int64_t>(TS, n - t * TS);
          for (int j = 0; j < cnt; ++j) {
            if (tcomp[j] == mycomp) continue;
            const float cj = tc[j];
            if (cj > bound) continue;
            const float dd = dsq_seq<D>(q, tp + j * nd, nd);
            const float mr = fmaxf(fmaxf(dd, cq), cj);
            const int32_t oj = tid_[j];
            const int32_t a = min(myid, oj), b = max(myid, oj);
            if (better(mr, a, b, bst)) {
              bst = Best{mr, a, b};
              bound = fminf(bound, mr);
            }
          }
        }
      } else {
        // Transposed: lane j holds the tile's point j; each query that needs the tile
        // takes the arg-min of its edges to the tile under the (w, a, b) order across the
        // wavefront. Edges above the query's bound are left out, as in the lane scan.
        float cx[MAXD];
#pragma unroll
        for (int c = 0; c < MAXD; ++c) cx[c] = (cv && c < nd) ? es[pt * nd + c] : 0.0f;
        const float cj = cv ? core[pt] : __builtin_huge_valf();
        const int32_t ccomp = cv ? comp[pt] : -1;
        const int32_t cid = cv ? (int32_t)perm[pt] : 0;
        for (unsigned long long need = need0; need != 0; need &= need - 1) {
          const int ql = __ffsll((long long)need) - 1;
          float qv[MAXD];
#pragma unroll
          for (int c = 0; c < MAXD; ++c) qv[c] = c < nd ? rdlane(q[c], ql) : 0.0f;
          const float cqq = rdlane(cq, ql), bq = rdlane(bound, ql);
          const int32_t compq = __builtin_amdgcn_readlane(mycomp, ql);
          const int32_t idq = __builtin_amdgcn_readlane(myid, ql);
          const bool el = cv && ccomp != compq && !(cj > bq);
          const float mr = el ? fmaxf(fmaxf(dsq_seq<D>(qv, cx, nd), cqq), cj) : __builtin_huge_valf();
          const float wm = wave_min(el ? mr : __builtin_huge_valf());
          const bool at = el && mr == wm;
          if (__ballot(at) == 0) continue;
          const int32_t a = min(idq, cid), b = max(idq, cid);
          const unsigned long long key = wave_min_u64(
              at ? (((unsigned long long)(uint32_t)a << 32) | (uint32_t)b) : ~0ull);
          const int32_t ka = (int32_t)(key >> 32), kb = (int32_t)(uint32_t)key;
          if (lane == ql && better(wm, ka, kb, bst)) bst = Best{wm, ka, kb};
        }
      }
      // only an edge below the bound read at this node can lower it
      publish_bound(comp_w, valid, mycomp, (!prune && bst.w < cw) ? bst.w : __builtin_huge_valf());
      continue;
    }
    // children: skip those no lane needs, walk the nearer one first
    const bool alive = live && !(cq > bound);
    const int32_t c0 = ncomp[2 * v], c1 = ncomp[2 * v + 1];
    const bool o0 = alive && c0 != -2 && !(c0 >= 0 && c0 == mycomp);
    const bool o1 = alive && c1 != -2 && !(c1 >= 0 && c1 == mycomp);
    const float d0 = o0 ? box_dist2<D>(q, nlo + 2 * v * nd, nhi + 2 * v * nd, nd)
                        : __builtin_huge_valf();
    const float d1 = o1 ? box_dist2<D>(q, nlo + (2 * v + 1) * nd, nhi + (2 * v + 1) * nd, nd)
                        : __builtin_huge_valf();
    const bool n0 = !__all(!o0 || d0 > bound), n1 = !__all(!o1 || d1 > bound);
    const float m0 = wave_min(d0), m1 = wave_min(d1);
    const int32_t near = (int32_t)(m1 < m0 ? 2 * v + 1 : 2 * v);
    const bool nn = m1 < m0 ? n1 : n0, nf = m1 < m0 ? n0 : n1;
    if (nf) stk[sp++] = near ^ 1;
    if (nn) stk[sp++] = near;
  }
  if (valid) {
    best_w[p] = bst.w;
    best_a[p] = bst.a;
    best_b[p] = bst.b;
  }
}

// Per-round bound seeding: each point tries the points of its own Morton tile and the
// two adjacent tiles, and every component's bound starts at the best such
// mutual-reachability edge to another component. These are real edges, so the bound is
// an upper bound of the component's minimum outgoing edge; the search then only has to
// certify within it (its pruning is strict, so edges equal to the bound are still seen)
// instead of starting from an unbounded radius.
template <int D>
__global__ __launch_bounds__(64) void boruvka_seed_kernel(
    const float* __restrict__ es, const float* __restrict__ core, const int32_t* __restrict__ comp,
    int64_t n, int dim, int64_t ntiles, unsigned* __restrict__ comp_w) {
  __shared__ float tp[3 * TS * MAXD];
  __shared__ float tc[3 * TS];
  __shared__ int32_t tcomp[3 * TS];
  const int lane = threadIdx.x;
  const int64_t qt = blockIdx.x;
  const int64_t p = qt * TS + lane;
  const bool valid = p < n;
  const int nd = D > 0 ? D : dim;
  for (int w = 0; w < 3; ++w) {
    const int64_t t = qt - 1 + w;
    const int64_t pt = t * TS + lane;
    const bool ok = t >= 0 && t < ntiles && pt < n;
    for (int c = 0; c < nd; ++c) tp[(w * TS + lane) * nd + c] = ok ? es[pt * nd + c] : 0.0f;
    tc[w * TS + lane] = ok ? core[pt] : __builtin_huge_valf();
    tcomp[w * TS + lane] = ok ? comp[pt] : -1;
  }
  __syncthreads();
  float q[MAXD];
#pragma unroll
  for (int c = 0; c < MAXD; ++c) q[c] = (valid && c < nd) ? es[p * nd + c] : 0.0f;
  const float cq = valid ? core[p] : 0.0f;
  const int32_t mycomp = valid ? comp[p] : -1;
  float best = __builtin_huge_valf();
  if (valid) {
    for (int j = 0; j < 3 * TS; ++j) {
      const int32_t cj = tcomp[j];
      if (cj < 0 || cj == mycomp) continue;
      const float dd = dsq_seq<D>(q, tp + j * nd, nd);
      const float mr = fmaxf(fmaxf(dd, cq), tc[j]);
      best = fminf(best, mr);
    }
  }
  publish_bound(comp_w, valid, mycomp, best);
}

// First round, from each point's k nearest (positions in the sorted order): every listed
// q has d(p, q) <= core(p), so its mutual-reachability weight is max(core(p), core(q)).
// When some listed q has core(q) <= core(p), that weight is core(p), which no edge of p can
// undercut; unlisted points are farther than the k-th distance unless one ties it. So
// without such a tie, the best listed edge under (w, min id, max id) is p's best edge.
__global__ void boruvka_round1_kernel(int64_t n, int k, const float* __restrict__ core,
                                      const int32_t* __restrict__ knn_pos,
                                      const uint8_t* __restrict__ knn_tie,
                                      const uint32_t* __restrict__ perm, float* __restrict__ pre_w,
                                      int32_t* __restrict__ pre_a, int32_t* __restrict__ pre_b,
                                      uint8_t* __restrict__ pre_done) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float cp = core[p];
  const int32_t me = (int32_t)perm[p];
  Best b{__builtin_huge_valf(), INT32_MAX, INT32_MAX};
  for (int m = 0; m < k; ++m) {
    const int32_t qpos = knn_pos[p * k + m];
    if (qpos < 0) continue;
    const float w = fmaxf(cp, core[qpos]);
    const int32_t o = (int32_t)perm[qpos];
    const int32_t a = min(me, o), bb = max(me, o);
    if (better(w, a, bb, b)) b = Best{w, a, bb};
  }
  pre_w[p] = b.w;
  pre_a[p] = b.a;
  pre_b[p] = b.b;
  pre_done[p] = (b.w == cp && !knn_tie[p]) ? 1 : 0;
}

// Later rounds: a point's best edge from the previous round is still a real edge out
// of its (merged) component when the far end lies in another component now, so it
// seeds that component's bound too.
__global__ void boruvka_prev_seed_kernel(int64_t n, const int32_t* __restrict__ comp,
                                         const uint32_t* __restrict__ perm,
                                         const int32_t* __restrict__ inv,
                                         const float* __restrict__ best_w,
                                         const int32_t* __restrict__ best_a,
                                         const int32_t* __restrict__ best_b,
                                         unsigned* __restrict__ comp_w) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float w = best_w[p];
  if (!(w < __builtin_huge_valf())) return;
  const int32_t me = (int32_t)perm[p];
  const int32_t other = best_a[p] == me ? best_b[p] : best_a[p];
  const int32_t c = comp[p];
  if (comp[inv[other]] != c) atomicMin(&comp_w[c], __float_as_uint(w));
}

__global__ void boruvka_select_kernel(int64_t n, const int32_t* __restrict__ comp,
                                      const unsigned* __restrict__ comp_w,
                                      const float* __restrict__ best_w,
                                      const int32_t* __restrict__ best_a,
                                      const int32_t* __restrict__ best_b,
                                      unsigned long long* __restrict__ comp_key) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float w = best_w[p];
  if (!(w < __builtin_huge_valf())) return;
  const int32_t c = comp[p];
  if (__float_as_uint(w) != comp_w[c]) return;
  const unsigned long long key = ((unsigned long long)(uint32_t)best_a[p] << 32) | (uint32_t)best_b[p];
  atomicMin(&comp_key[c], key);
}

__global__ void tile_comp_kernel(const int32_t* __restrict__ comp, int64_t n, int64_t ntiles,
                                 int32_t* __restrict__ tile_comp) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t t = (int64_t)blockIdx.x * 4 + wave;
  if (t >= ntiles) return;
  const int64_t p = t * TS + lane;
  const int32_t c = p < n ? comp[p] : comp[t * TS];
  const int32_t c0 = __shfl(c, 0, 64);
  const bool same = __all(c == c0);
  if (lane == 0) tile_comp[t] = same ? c0 : -1;
}

struct Sorted {
  Buf<float> es;
  Buf<uint32_t> perm;
  Buf<float> blo, bhi;  // tile boxes
  Buf<float> nlo, nhi;  // tree node boxes, nodes 1 .. 2L-1 (leaves L .. 2L-1)
  int64_t ntiles = 0, L = 1;
};

void kd_sort(const float* d_emb, int64_t n, int dim, Sorted& S, hipStream_t s, DevPool& pool) {
  S.ntiles = ceil_div(n, TS);
  S.L = 1;
  int h = 0;
  while (S.L < S.ntiles) {
    S.L <<= 1;
    ++h;
  }
  S.perm.alloc(pool, n);
  const unsigned g = (unsigned)ceil_div(n, 256);
  iota_kernel<<<g, 256, 0, s>>>(n, S.perm.p);
  {
    Buf<float> tlo(pool, S.ntiles * dim), thi(pool, S.ntiles * dim);
    Buf<float> olo(pool, (S.L / 2) * dim + dim), ohi(pool, (S.L / 2) * dim + dim);
    Buf<uint64_t> key(pool, n);
    for (int lev = 0; lev < h; ++lev) {
      perm_tile_box_kernel<<<(unsigned)ceil_div(S.ntiles, 4), 256, 0, s>>>(
          d_emb, S.perm.p, n, dim, S.ntiles, tlo.p, thi.p);
      const int64_t span = S.L >> lev;
      const int64_t nodes = ceil_div(S.ntiles, span);
      kd_node_box_kernel<<<(unsigned)nodes, 256, 0, s>>>(tlo.p, thi.p, S.ntiles, span, dim,
                                                         olo.p, ohi.p);
      kd_key_kernel<<<g, 256, 0, s>>>(d_emb, S.perm.p, n, dim, h - lev, olo.p, ohi.p, key.p,
                                      S.perm.p);
      radix_sort_pairs(key.p, S.perm.p, n, lev + 16, s, pool);
    }
  }
  S.es.alloc(pool, n * dim);
  gather_kernel<<<(unsigned)ceil_div(n * dim, 256), 256, 0, s>>>(d_emb, S.perm.p, n, dim, S.es.p);
  S.blo.alloc(pool, S.ntiles * dim);
  S.bhi.alloc(pool, S.ntiles * dim);
  tile_box_kernel<<<(unsigned)ceil_div(S.ntiles, 4), 256, 0, s>>>(S.es.p, n, dim, S.ntiles,
                                                                  S.blo.p, S.bhi.p);
  S.nlo.alloc(pool, 2 * S.L * dim);
  S.nhi.alloc(pool, 2 * S.L * dim);
  bvh_leaf_kernel<<<(unsigned)ceil_div(S.L, 256), 256, 0, s>>>(S.blo.p, S.bhi.p, S.ntiles, S.L,
                                                               dim, S.nlo.p, S.nhi.p);
  for (int64_t lev = S.L / 2; lev >= 1; lev /= 2)
    bvh_merge_kernel<<<(unsigned)ceil_div(lev, 256), 256, 0, s>>>(lev, dim, S.nlo.p, S.nhi.p);
  EVOC_HIP(hipGetLastError());
}

// core distances in sorted order; with knn_pos / knn_tie (k <= 8 only: returns false
// otherwise) also each point's k nearest and the tie flag
bool core_sorted(const Sorted& S, int64_t n, int dim, int min_samples, float* d_core_sorted,
                 hipStream_t s, int32_t* knn_pos = nullptr, uint8_t* knn_tie = nullptr) {
  const int k = (int)std::min<int64_t>(min_samples, n - 1);
  EVOC_CHECK(k <= KCORE_MAX, "min_samples must be <= 64");
  if (k <= 0) {
    EVOC_HIP(hipMemsetAsync(d_core_sorted, 0, n * sizeof(float), s));
    return false;
  }
  const unsigned g = (unsigned)S.ntiles;
  const bool list = knn_pos && knn_tie && k <= 8;
  if (list) {
    if (dim == 8)
      core_kernel<8, 8, true><<<g, 64, 0, s>>>(S.es.p, n, dim, S.ntiles, S.L, S.nlo.p, S.nhi.p,
                                               k, d_core_sorted, knn_pos, knn_tie);
    else
      core_kernel<0, 8, true><<<g, 64, 0, s>>>(S.es.p, n, dim, S.ntiles, S.L, S.nlo.p, S.nhi.p,
                                               k, d_core_sorted, knn_pos, knn_tie);
  } else if (dim == 8 && k <= 8) {
    core_kernel<8, 8, false><<<g, 64, 0, s>>>(S.es.p, n, dim, S.ntiles, S.L, S.nlo.p, S.nhi.p,
                                              k, d_core_sorted, nullptr, nullptr);
  } else if (dim == 8 && k <= KCORE) {
    core_kernel<8, KCORE, false><<<g, 64, 0, s>>>(S.es.p, n, dim, S.ntiles, S.L, S.nlo.p,
                                                  S.nhi.p, k, d_core_sorted, nullptr, nullptr);
  } else if (k <= 8) {
    core_kernel<0, 8, false><<<g, 64, 0, s>>>(S.es.p, n, dim, S.ntiles, S.L, S.nlo.p, S.nhi.p,
                                              k, d_core_sorted, nullptr, nullptr);
  } else if (k > KCORE) {
    core_kernel<0, KCORE_MAX, false><<<g, 64, 0, s>>>(S.es.p, n, dim, S.ntiles, S.L, S.nlo.p,
                                                      S.nhi.p, k, d_core_sorted, nullptr, nullptr);
  } else {
    core_kernel<0, KCORE, false><<<g, 64, 0, s>>>(S.es.p, n, dim, S.ntiles, S.L, S.nlo.p,
                                                  S.nhi.p, k, d_core_sorted, nullptr, nullptr);
  }
  EVOC_HIP(hipGetLastError());
  return list;
}

__global__ void unpermute_kernel(const float* __restrict__ src, const uint32_t* __restrict__ perm,
                                 int64_t n, float* __restrict__ dst) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) dst[perm[p]] = src[p];
}

// ---------------------------------------------------------------- Boruvka merge step
// Each component's chosen edge hooks it onto the component at the other end. Under a
// strict total order on edges the hooks form a forest whose only cycles are mutual
// pairs (two components choosing the same edge); a pair is rooted at its smaller id,
// and its edge is emitted once, by that root.
__global__ void inv_perm_kernel(const uint32_t* __restrict__ perm, int64_t n,
                                int32_t* __restrict__ inv) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) inv[perm[p]] = (int32_t)p;
}

__global__ void hook_kernel(int64_t ncomp, const unsigned long long* __restrict__ comp_key,
                            const int32_t* __restrict__ comp, const int32_t* __restrict__ inv,
                            int32_t* __restrict__ partner) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncomp) return;
  const unsigned long long key = comp_key[c];
  if (key == ~0ull) {
    partner[c] = (int32_t)c;
    return;
  }
  const int32_t ca = comp[inv[(int32_t)(key >> 32)]];
  const int32_t cb = comp[inv[(int32_t)(key & 0xffffffffu)]];
  partner[c] = ca == (int32_t)c ? cb : ca;
}

__global__ void link_kernel(int64_t ncomp, const int32_t* __restrict__ partner,
                            const unsigned long long* __restrict__ comp_key,
                            const unsigned* __restrict__ comp_w, int64_t cap,
                            int32_t* __restrict__ parent, unsigned* __restrict__ nedge,
                            float* __restrict__ ew2, int32_t* __restrict__ ea,
                            int32_t* __restrict__ eb) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncomp) return;
  const int32_t o = partner[c];
  if (o == (int32_t)c) {
    parent[c] = o;
    return;
  }
  const bool mutual = partner[o] == (int32_t)c;
  parent[c] = (mutual && (int32_t)c < o) ? (int32_t)c : o;
  if (mutual && (int32_t)c > o) return;
  const unsigned idx = atomicAdd(nedge, 1u);
  if (idx < cap) {
    const unsigned long long key = comp_key[c];
    ew2[idx] = __uint_as_float(comp_w[c]);
    ea[idx] = (int32_t)(key >> 32);
    eb[idx] = (int32_t)(key & 0xffffffffu);
  }
}

// pointer jumping toward the roots; `busy` is set while some pointer still moves
__global__ void jump_kernel(int64_t ncomp, int32_t* __restrict__ parent,
                            unsigned* __restrict__ busy) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncomp) return;
  int32_t pc = parent[c];
  for (int it = 0; it < 32; ++it) {
    const int32_t pp = parent[pc];
    if (pp == pc) return;
    pc = pp;
    parent[c] = pc;
  }
  if (parent[pc] != pc) *busy = 1u;
}

__global__ void root_flag_kernel(int64_t ncomp, const int32_t* __restrict__ parent,
                                 int64_t* __restrict__ flag) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c <= ncomp) flag[c] = (c < ncomp && parent[c] == (int32_t)c) ? 1 : 0;
}

__global__ void relabel_kernel(int64_t n, const int32_t* __restrict__ parent,
                               const int64_t* __restrict__ rank, int32_t* __restrict__ comp) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) comp[p] = (int32_t)rank[parent[comp[p]]];
}

// final order of the edges: (sqrt(w^2), u, v) ascending, by two stable radix sorts
__global__ void edge_ab_key_kernel(int64_t m, const int32_t* __restrict__ ea,
                                   const int32_t* __restrict__ eb, uint64_t* __restrict__ key,
                                   uint32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  key[i] = ((uint64_t)(uint32_t)ea[i] << 32) | (uint32_t)eb[i];
  idx[i] = (uint32_t)i;
}
__global__ void edge_w_key_kernel(int64_t m, const float* __restrict__ ew2,
                                  const uint32_t* __restrict__ idx, uint64_t* __restrict__ key) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) key[i] = __float_as_uint(sqrtf(ew2[idx[i]]));
}
__global__ void edge_gather_kernel(int64_t m, const float* __restrict__ ew2,
                                   const int32_t* __restrict__ ea, const int32_t* __restrict__ eb,
                                   const uint32_t* __restrict__ idx, int32_t* __restrict__ u,
                                   int32_t* __restrict__ v, float* __restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t j = idx[i];
  u[i] = ea[j];
  v[i] = eb[j];
  w[i] = sqrtf(ew2[j]);
}

}  // namespace

void core_sq_device(const float* d_emb, int64_t n, int dim, int min_samples, float* d_core,
                    hipStream_t s, DevPool& pool) {
  EVOC_CHECK(dim >= 1 && dim <= MAXD, "core distances: dim must be in [1, 16]");
  if (n <= 0) return;
  if (n == 1) {
    EVOC_HIP(hipMemsetAsync(d_core, 0, sizeof(float), s));
    return;
  }
  Sorted S;
  kd_sort(d_emb, n, dim, S, s, pool);
  Buf<float> cs(pool, n);
  core_sorted(S, n, dim, min_samples, cs.p, s);
  unpermute_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(cs.p, S.perm.p, n, d_core);
  EVOC_HIP(hipGetLastError());
}

void mst_device(const float* d_emb, int64_t n, int dim, int min_samples, int32_t* h_u,
                int32_t* h_v, float* h_w, hipStream_t s, DevPool& pool, int* rounds) {
  EVOC_CHECK(dim >= 1 && dim <= MAXD, "mst: dim must be in [1, 16]");
  if (rounds) *rounds = 0;
  if (n <= 1) return;
  Sorted S;
  kd_sort(d_emb, n, dim, S, s, pool);
  Buf<float> core(pool, n);
  const int kc = (int)std::min<int64_t>(min_samples, n - 1);
  Buf<int32_t> knn_pos(pool, n * std::max(kc, 1));
  Buf<uint8_t> knn_tie(pool, n);
  const bool have_list = core_sorted(S, n, dim, min_samples, core.p, s, knn_pos.p, knn_tie.p);
  Buf<float> pre_w;
  Buf<int32_t> pre_a, pre_b;
  Buf<uint8_t> pre_done;
  if (have_list) {
    pre_w.alloc(pool, n);
    pre_a.alloc(pool, n);
    pre_b.alloc(pool, n);
    pre_done.alloc(pool, n);
    boruvka_round1_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(
        n, kc, core.p, knn_pos.p, knn_tie.p, S.perm.p, pre_w.p, pre_a.p, pre_b.p, pre_done.p);
  }
  Buf<int32_t> comp(pool, n), tile_comp(pool, S.ntiles), ba(pool, n), bb(pool, n);
  Buf<int32_t> ncomp_node(pool, 2 * S.L), inv(pool, n), partner(pool, n), parent(pool, n);
  Buf<float> bw(pool, n), ew2(pool, n);
  Buf<int32_t> ea(pool, n), eb(pool, n);
  Buf<unsigned> comp_w(pool, n), ctr(pool, 2);  // ctr: edge count, jump flag
  Buf<unsigned long long> comp_key(pool, n);
  Buf<int64_t> flag(pool, n + 1), rank(pool, n + 1);
  const unsigned gn = (unsigned)ceil_div(n, 256);
  iota_kernel<<<gn, 256, 0, s>>>(n, reinterpret_cast<uint32_t*>(comp.p));  // own components
  inv_perm_kernel<<<gn, 256, 0, s>>>(S.perm.p, n, inv.p);
  EVOC_HIP(hipMemsetAsync(ctr.p, 0, 2 * sizeof(unsigned), s));
  int64_t ncomp = n;
  unsigned nedge = 0;
  int r = 0;
  while (ncomp > 1) {
    ++r;
    tile_comp_kernel<<<(unsigned)ceil_div(S.ntiles, 4), 256, 0, s>>>(comp.p, n, S.ntiles,
                                                                     tile_comp.p);
    bvh_comp_leaf_kernel<<<(unsigned)ceil_div(S.L, 256), 256, 0, s>>>(tile_comp.p, S.ntiles, S.L,
                                                                      ncomp_node.p);
    for (int64_t lev = S.L / 2; lev >= 1; lev /= 2)
      bvh_comp_merge_kernel<<<(unsigned)ceil_div(lev, 256), 256, 0, s>>>(lev, ncomp_node.p);
    EVOC_HIP(hipMemsetAsync(comp_w.p, 0x7f, ncomp * sizeof(unsigned), s));  // 0x7f7f7f7f > any finite w
    EVOC_HIP(hipMemsetAsync(comp_key.p, 0xff, ncomp * sizeof(unsigned long long), s));
    const unsigned g = (unsigned)S.ntiles;
    if (r > 1)
      boruvka_prev_seed_kernel<<<gn, 256, 0, s>>>(n, comp.p, S.perm.p, inv.p, bw.p, ba.p, bb.p,
                                                  comp_w.p);
    if (dim == 8)
      boruvka_seed_kernel<8><<<g, 64, 0, s>>>(S.es.p, core.p, comp.p, n, dim, S.ntiles, comp_w.p);
    else
      boruvka_seed_kernel<0><<<g, 64, 0, s>>>(S.es.p, core.p, comp.p, n, dim, S.ntiles, comp_w.p);
    const bool first = r == 1 && have_list;
    const float* pw = first ? pre_w.p : nullptr;
    const int32_t* pa = first ? pre_a.p : nullptr;
    const int32_t* pbb = first ? pre_b.p : nullptr;
    const uint8_t* pd = first ? pre_done.p : nullptr;
    if (dim == 8)
      boruvka_search_kernel<8><<<g, 64, 0, s>>>(S.es.p, core.p, comp.p, S.perm.p, n, dim, S.L,
                                                S.nlo.p, S.nhi.p, ncomp_node.p, comp_w.p, bw.p,
                                                ba.p, bb.p, pw, pa, pbb, pd);
    else
      boruvka_search_kernel<0><<<g, 64, 0, s>>>(S.es.p, core.p, comp.p, S.perm.p, n, dim, S.L,
                                                S.nlo.p, S.nhi.p, ncomp_node.p, comp_w.p, bw.p,
                                                ba.p, bb.p, pw, pa, pbb, pd);
    boruvka_select_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(n, comp.p, comp_w.p, bw.p,
                                                                     ba.p, bb.p, comp_key.p);
    EVOC_HIP(hipGetLastError());
    const unsigned gc = (unsigned)ceil_div(ncomp, 256);
    hook_kernel<<<gc, 256, 0, s>>>(ncomp, comp_key.p, comp.p, inv.p, partner.p);
    link_kernel<<<gc, 256, 0, s>>>(ncomp, partner.p, comp_key.p, comp_w.p, n - 1, parent.p,
                                   ctr.p, ew2.p, ea.p, eb.p);
    unsigned h[2] = {0, 1};
    for (int it = 0; h[1]; ++it) {
      // 32 jumps per launch per pointer: a forest of any depth below 2^31 settles in a
      // few launches; a cycle (which a strict edge order rules out) would never settle
      EVOC_CHECK(it < 64, "mst: Boruvka hooks do not form a forest");
      EVOC_HIP(hipMemsetAsync(ctr.p + 1, 0, sizeof(unsigned), s));
      jump_kernel<<<gc, 256, 0, s>>>(ncomp, parent.p, ctr.p + 1);
      EVOC_HIP(hipMemcpyAsync(h, ctr.p, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
      EVOC_HIP(hipStreamSynchronize(s));
    }
    EVOC_CHECK(h[0] <= (unsigned)(n - 1), "mst: more Boruvka edges than n - 1");
    if (h[0] == nedge) break;  // no component has an outgoing edge left
    nedge = h[0];
    root_flag_kernel<<<(unsigned)ceil_div(ncomp + 1, 256), 256, 0, s>>>(ncomp, parent.p, flag.p);
    exclusive_scan_i64(flag.p, rank.p, ncomp + 1, s, pool);
    relabel_kernel<<<gn, 256, 0, s>>>(n, parent.p, rank.p, comp.p);
    int64_t nc = 0;
    EVOC_HIP(hipMemcpyAsync(&nc, rank.p + ncomp, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipStreamSynchronize(s));
    ncomp = nc;
  }
  if (rounds) *rounds = r;
  const int64_t m = nedge;
  if (m > 0) {
    Buf<uint64_t> key(pool, m);
    Buf<uint32_t> idx(pool, m);
    const unsigned gm = (unsigned)ceil_div(m, 256);
    edge_ab_key_kernel<<<gm, 256, 0, s>>>(m, ea.p, eb.p, key.p, idx.p);
    radix_sort_pairs(key.p, idx.p, m, 64, s, pool);
    edge_w_key_kernel<<<gm, 256, 0, s>>>(m, ew2.p, idx.p, key.p);
    radix_sort_pairs(key.p, idx.p, m, 32, s, pool);
    edge_gather_kernel<<<gm, 256, 0, s>>>(m, ew2.p, ea.p, eb.p, idx.p, ba.p, bb.p, bw.p);
    EVOC_HIP(hipGetLastError());
    EVOC_HIP(hipMemcpyAsync(h_u, ba.p, m * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(h_v, bb.p, m * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(h_w, bw.p, m * sizeof(float), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipStreamSynchronize(s));
  }
  for (int64_t i = m; i < n - 1; ++i) {
    h_u[i] = -1;
    h_v[i] = -1;
    h_w[i] = __builtin_huge_valf();
  }
}

}  // namespace evoc
