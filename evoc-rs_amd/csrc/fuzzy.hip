// Fuzzy simplicial set (build_fuzzy_simplicial_set, src/graph/fuzzy_graph.rs:275-295).
//
//  sigma:   one row per thread: rho = d[0], 64-step bisection on sum(exp(-(d-rho)/s)),
//           f32 sums compared in f64 (smooth_knn_dist, fuzzy_graph.rs:33-98), then the
//           directed weights exp(-max(0, d-rho)/sigma) (knn_to_coo, :114-159).
//  symmetrise (symmetrise_graph, :174-228): every directed edge i->j is written into
//           row i as (j, fwd) and into row j as (i, bwd) (in-degree counted, scanned,
//           scattered); each row is then ordered by (column, direction) with a rank
//           sort and adjacent fwd/bwd pairs merge into w_ij + w_ji - w_ij*w_ji. Rows
//           come out sorted by column, so the CSR is canonical whatever the scatter
//           order was.
#include "kernels.h"

namespace evoc {
namespace {

constexpr int KMAXF = 64;  // the row's distances in registers: 32 or 64 of them

// canonical f32 exponential: exp evaluated in f64 and rounded once to f32 (correctly
// rounded except in vanishingly rare double-rounding cases), identical on host and device
__device__ __forceinline__ float cexp(float x) { return (float)exp((double)x); }

template <int KMAXF>
__global__ __launch_bounds__(256) void sigma_kernel(const float* __restrict__ dist,
                                                    const int32_t* __restrict__ idx, int64_t n,
                                                    int k, double target, float* __restrict__ sigma,
                                                    float* __restrict__ rho,
                                                    float* __restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float d[KMAXF];
#pragma unroll
  for (int t = 0; t < KMAXF; ++t) d[t] = t < k ? dist[i * k + t] : 0.0f;
  const float r = d[0];
  const float fmax_ = 3.40282347e+38f;
  float lo = 0.0f, hi = fmax_, mid = 1.0f;
  for (int it = 0; it < 64; ++it) {
    float val = 0.0f;
#pragma unroll
    for (int t = 0; t < KMAXF; ++t)
      if (t < k) {
        const float adjusted = fmaxf(d[t] - r, 0.0f);
        val += cexp(-(adjusted / mid));
      }
    if (fabs((double)val - target) < 1e-5) break;
    if ((double)val > target) {
      hi = mid;
      mid = (lo + hi) / 2.0f;
    } else {
      lo = mid;
      if (hi == fmax_) mid = mid * 2.0f;
      else mid = (lo + hi) / 2.0f;
    }
  }
  sigma[i] = mid;
  rho[i] = r;
#pragma unroll
  for (int t = 0; t < KMAXF; ++t)
    if (t < k) {
      const float adjusted = fmaxf(d[t] - r, 0.0f);
      float wt;
      if (mid > 0.0f) wt = cexp(-(adjusted / mid));
      else if (adjusted > 0.0f) wt = 0.0f;
      else wt = 1.0f;
      w[i * k + t] = wt;
    }
}

__device__ __forceinline__ bool edge_valid(int32_t j, int64_t i, int64_t n) {
  return j >= 0 && (int64_t)j != i && (int64_t)j < n;
}

__global__ void degree_kernel(const int32_t* __restrict__ idx, int64_t n, int k,
                              int64_t* __restrict__ outdeg, int64_t* __restrict__ rowcap,
                              unsigned* __restrict__ indeg, int symmetrise) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t c = 0;
  for (int t = 0; t < k; ++t) {
    const int32_t j = idx[i * k + t];
    if (edge_valid(j, i, n)) {
      ++c;
      if (symmetrise) atomicAdd(&indeg[j], 1u);
    }
  }
  outdeg[i] = c;
  rowcap[i] = c;  // the in-degree part is added by add_indeg_kernel
}

__global__ void add_indeg_kernel(int64_t n, const unsigned* __restrict__ indeg,
                                 int64_t* __restrict__ rowcap) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rowcap[i] += indeg[i];
}

// records: key = (col << 1) | dir (dir 0 = i->col, 1 = col->i), value = weight
__global__ void fill_kernel(const int32_t* __restrict__ idx, const float* __restrict__ w,
                            int64_t n, int k, const int64_t* __restrict__ off,
                            const int64_t* __restrict__ outdeg, unsigned* __restrict__ cursor,
                            uint64_t* __restrict__ rkey, float* __restrict__ rval,
                            int symmetrise) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t p = off[i];
  for (int t = 0; t < k; ++t) {
    const int32_t j = idx[i * k + t];
    if (!edge_valid(j, i, n)) continue;
    const float wt = w[i * k + t];
    rkey[p] = ((uint64_t)(uint32_t)j << 1);
    rval[p] = wt;
    ++p;
    if (symmetrise) {
      const unsigned c = atomicAdd(&cursor[j], 1u);
      const int64_t q = off[j] + outdeg[j] + c;
      rkey[q] = ((uint64_t)(uint64_t)i << 1) | 1u;
      rval[q] = wt;
    }
  }
}

__device__ __forceinline__ float fuzzy_union(float wij, float wji) {
  const float s = wij + wji;
  const float p = wij * wji;
  return s - p;
}

// one wavefront per row with at most 64 records; longer rows are queued
__global__ __launch_bounds__(256) void row_merge_small(
    int64_t n, const int64_t* __restrict__ off, const int64_t* __restrict__ rowcap,
    const uint64_t* __restrict__ rkey, const float* __restrict__ rval,
    int32_t* __restrict__ tcol, float* __restrict__ tw, int64_t* __restrict__ cnt,
    int* __restrict__ nbig, int64_t* __restrict__ big) {
  __shared__ uint64_t sk[4][64];
  __shared__ float sv[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + wave;
  // no early return: every wave reaches the barrier below
  const int64_t L = i < n ? rowcap[i] : 0;
  const bool small = i < n && L <= 64;
  if (i < n && L > 64 && lane == 0) {
    const int pos = atomicAdd(nbig, 1);
    big[pos] = i;
  }
  const int64_t o = i < n ? off[i] : 0;
  const bool have = small && lane < L;
  const uint64_t key = have ? rkey[o + lane] : ~0ull;
  const float val = have ? rval[o + lane] : 0.0f;
  int rank = 0;
  const int Lw = small ? (int)L : 0;
  for (int m = 0; m < Lw; ++m) {
    const uint64_t km = __shfl(key, m, 64);
    rank += km < key;
  }
  if (have) {
    sk[wave][rank] = key;
    sv[wave][rank] = val;
  }
  __syncthreads();
  bool head = false;
  float ws = 0.0f;
  int32_t col = 0;
  if (have) {
    const uint64_t kp = sk[wave][lane];
    col = (int32_t)(kp >> 1);
    const float wp = sv[wave][lane];
    if ((kp & 1) == 0) {
      float wji = 0.0f;
      if (lane + 1 < L && (sk[wave][lane + 1] >> 1) == (kp >> 1)) wji = sv[wave][lane + 1];
      ws = fuzzy_union(wp, wji);
      head = true;
    } else {
      const bool paired = lane > 0 && (sk[wave][lane - 1] >> 1) == (kp >> 1);
      if (!paired) {
        ws = fuzzy_union(0.0f, wp);
        head = true;
      }
    }
    head = head && ws > 0.0f;
  }
  const unsigned long long hb = __ballot(head);
  if (head) {
    const int pos = __popcll(hb & ((1ull << lane) - 1ull));
    tcol[o + pos] = col;
    tw[o + pos] = ws;
  }
  if (small && lane == 0) cnt[i] = __popcll(hb);
}

// long rows: one workgroup each, rank sort through global scratch, serial compaction
__global__ __launch_bounds__(256) void row_merge_big(
    const int* __restrict__ nbig, const int64_t* __restrict__ big,
    const int64_t* __restrict__ off, const int64_t* __restrict__ rowcap,
    const uint64_t* __restrict__ rkey, const float* __restrict__ rval,
    uint64_t* __restrict__ skey, float* __restrict__ sval, int32_t* __restrict__ tcol,
    float* __restrict__ tw, int64_t* __restrict__ cnt) {
  const int nb = *nbig;
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const int64_t i = big[b];
    const int64_t o = off[i], L = rowcap[i];
    for (int64_t e = threadIdx.x; e < L; e += blockDim.x) {
      const uint64_t ke = rkey[o + e];
      int64_t r = 0;
      for (int64_t m = 0; m < L; ++m) r += rkey[o + m] < ke;
      skey[o + r] = ke;
      sval[o + r] = rval[o + e];
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t c = 0;
      for (int64_t p = 0; p < L; ++p) {
        const uint64_t kp = skey[o + p];
        const float wp = sval[o + p];
        float ws;
        if ((kp & 1) == 0) {
          float wji = 0.0f;
          if (p + 1 < L && (skey[o + p + 1] >> 1) == (kp >> 1)) wji = sval[o + p + 1];
          ws = fuzzy_union(wp, wji);
        } else {
          if (p > 0 && (skey[o + p - 1] >> 1) == (kp >> 1)) continue;
          ws = fuzzy_union(0.0f, wp);
        }
        if (ws > 0.0f) {
          tcol[o + c] = (int32_t)(kp >> 1);
          tw[o + c] = ws;
          ++c;
        }
      }
      cnt[i] = c;
    }
    __syncthreads();
  }
}

// directed graph (no symmetrisation): rows keep kNN order
__global__ void directed_kernel(int64_t n, const int64_t* __restrict__ off,
                                const int64_t* __restrict__ rowcap,
                                const uint64_t* __restrict__ rkey, const float* __restrict__ rval,
                                int32_t* __restrict__ tcol, float* __restrict__ tw,
                                int64_t* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t o = off[i];
  for (int64_t p = 0; p < rowcap[i]; ++p) {
    tcol[o + p] = (int32_t)(rkey[o + p] >> 1);
    tw[o + p] = rval[o + p];
  }
  cnt[i] = rowcap[i];
}

__global__ void compact_kernel(int64_t n, const int64_t* __restrict__ off,
                               const int64_t* __restrict__ indptr, const int64_t* __restrict__ cnt,
                               const int32_t* __restrict__ tcol, const float* __restrict__ tw,
                               int32_t* __restrict__ col, float* __restrict__ wout) {
  // one wavefront per row
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + wave;
  if (i >= n) return;
  const int64_t src = off[i], dst = indptr[i], c = cnt[i];
  for (int64_t p = lane; p < c; p += 64) {
    col[dst + p] = tcol[src + p];
    wout[dst + p] = tw[src + p];
  }
}

__global__ void set_last(int64_t* indptr, int64_t n, const int64_t* cnt) {
  indptr[n] = indptr[n - 1] + cnt[n - 1];
}

}  // namespace

void fuzzy_device(const int32_t* d_idx, const float* d_dist, int64_t n, int k, float effective_k,
                  bool symmetrise, float* d_sigma, float* d_rho, Graph& out, hipStream_t s,
                  DevPool& pool) {
  EVOC_CHECK(k >= 1 && k <= KMAXF, "fuzzy: k must be in [1, 64]");
  EVOC_CHECK(n >= 1, "fuzzy: empty graph");
  // effective_n_neighbors.to_usize().unwrap_or(k) (fuzzy_graph.rs:284-286)
  const size_t kt = (effective_k > -1.0f && std::isfinite(effective_k)) ? (size_t)effective_k : (size_t)k;
  const double target = std::log2((double)kt);
  Buf<float> w(pool, n * k);
  Buf<float> sig, rh;
  if (!d_sigma) { sig.alloc(pool, n); d_sigma = sig.p; }
  if (!d_rho) { rh.alloc(pool, n); d_rho = rh.p; }
  const unsigned g256 = (unsigned)ceil_div(n, 256);
  (k <= 32 ? sigma_kernel<32> : sigma_kernel<64>)<<<g256, 256, 0, s>>>(d_dist, d_idx, n, k, target,
                                                                       d_sigma, d_rho, w.p);
  Buf<int64_t> outdeg(pool, n), rowcap(pool, n), off(pool, n), cnt(pool, n);
  Buf<unsigned> indeg(pool, n), cursor(pool, n);
  EVOC_HIP(hipMemsetAsync(indeg.p, 0, n * sizeof(unsigned), s));
  EVOC_HIP(hipMemsetAsync(cursor.p, 0, n * sizeof(unsigned), s));
  degree_kernel<<<g256, 256, 0, s>>>(d_idx, n, k, outdeg.p, rowcap.p, indeg.p, symmetrise ? 1 : 0);
  if (symmetrise) add_indeg_kernel<<<g256, 256, 0, s>>>(n, indeg.p, rowcap.p);
  exclusive_scan_i64(rowcap.p, off.p, n, s, pool);
  const int64_t cap = symmetrise ? 2 * n * k : n * k;
  Buf<uint64_t> rkey(pool, cap);
  Buf<float> rval(pool, cap);
  Buf<int32_t> tcol(pool, cap);
  Buf<float> tw(pool, cap);
  fill_kernel<<<g256, 256, 0, s>>>(d_idx, w.p, n, k, off.p, outdeg.p, cursor.p, rkey.p, rval.p,
                                   symmetrise ? 1 : 0);
  if (symmetrise) {
    Buf<int> nbig(pool, 1);
    Buf<int64_t> big(pool, n);
    EVOC_HIP(hipMemsetAsync(nbig.p, 0, sizeof(int), s));
    row_merge_small<<<(unsigned)ceil_div(n, 4), 256, 0, s>>>(n, off.p, rowcap.p, rkey.p, rval.p,
                                                            tcol.p, tw.p, cnt.p, nbig.p, big.p);
    Buf<uint64_t> skey(pool, cap);
    Buf<float> sval(pool, cap);
    row_merge_big<<<1024, 256, 0, s>>>(nbig.p, big.p, off.p, rowcap.p, rkey.p, rval.p, skey.p,
                                       sval.p, tcol.p, tw.p, cnt.p);
  } else {
    directed_kernel<<<g256, 256, 0, s>>>(n, off.p, rowcap.p, rkey.p, rval.p, tcol.p, tw.p, cnt.p);
  }
  out.indptr.alloc(pool, n + 1);
  exclusive_scan_i64(cnt.p, out.indptr.p, n, s, pool);
  set_last<<<1, 1, 0, s>>>(out.indptr.p, n, cnt.p);
  int64_t nnz = 0;
  EVOC_HIP(hipMemcpyAsync(&nnz, out.indptr.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
  out.n = n;
  out.nnz = nnz;
  out.indices.alloc(pool, nnz);
  out.weights.alloc(pool, nnz);
  compact_kernel<<<(unsigned)ceil_div(n, 4), 256, 0, s>>>(n, off.p, out.indptr.p, cnt.p, tcol.p,
                                                         tw.p, out.indices.p, out.weights.p);
  EVOC_HIP(hipGetLastError());
}

}  // namespace evoc
