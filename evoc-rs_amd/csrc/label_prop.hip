// Device parts of the label-propagation initial layout (label_propagation_init,
// src/graph/label_prop.rs:537-564).
//
// partition_means: the data coarsening of one level, (R^T_norm)_l1 X
// (src/utils/sparse.rs:479-501 row L1 normalisation, then the sparse-dense product).
// Row a of the result is coef[a] * sum of the member rows of partition a, accumulated
// member by member in ascending node order, each term rounded as coef * x before the
// add -- the order the sparse product walks a CSR row -- so the f32 values equal the
// host formulation bit for bit. One workgroup per (partition, 256 columns), one thread
// per column; member rows are loaded up to two batches of 32 ahead of the dependent adds.
#include "kernels.h"

namespace evoc {
namespace {

constexpr int PM_THREADS = 256;
constexpr int PM_AHEAD = 32;  // member rows loaded ahead of their in-order adds

// one workgroup per (partition, stripe of 256 columns)
__global__ __launch_bounds__(PM_THREADS) void partition_means_kernel(
    const float* __restrict__ X, int64_t ld, int64_t d, const int64_t* __restrict__ mptr,
    const int32_t* __restrict__ mem, const float* __restrict__ coef, float* __restrict__ out) {
  const int64_t a = blockIdx.x;
  const int64_t m0 = mptr[a], m1 = mptr[a + 1];
  const float cf = coef[a];
  {
    const int64_t c = (int64_t)blockIdx.y * PM_THREADS + threadIdx.x;
    if (c >= d) return;
    float acc = 0.0f;
    int64_t m = m0;
    // batches of PM_AHEAD member rows; the next batch's loads are issued before this
    // batch's adds, so two batches are in flight
    float v[PM_AHEAD];
    if (m + PM_AHEAD <= m1) {
#pragma unroll
      for (int u = 0; u < PM_AHEAD; ++u) v[u] = X[(int64_t)mem[m + u] * ld + c];
    }
    for (; m + PM_AHEAD <= m1; m += PM_AHEAD) {
      const bool more = m + 2 * PM_AHEAD <= m1;
      float w[PM_AHEAD];
      if (more) {
#pragma unroll
        for (int u = 0; u < PM_AHEAD; ++u) w[u] = X[(int64_t)mem[m + PM_AHEAD + u] * ld + c];
      }
#pragma unroll
      for (int u = 0; u < PM_AHEAD; ++u) {
        const float t = cf * v[u];
        acc = acc + t;
      }
      if (more) {
#pragma unroll
        for (int u = 0; u < PM_AHEAD; ++u) v[u] = w[u];
      }
    }
    for (; m < m1; ++m) {
      const float t = cf * X[(int64_t)mem[m] * ld + c];
      acc = acc + t;
    }
    out[a * d + c] = acc;
  }
}

}  // namespace

void partition_means_dev(const float* dX, int64_t ld, int64_t d, const int64_t* d_mptr,
                         const int32_t* d_mem, const float* d_coef, int64_t np, float* d_out,
                         hipStream_t s) {
  const dim3 grid((unsigned)np, (unsigned)ceil_div(d, PM_THREADS));
  partition_means_kernel<<<grid, PM_THREADS, 0, s>>>(dX, ld, d, d_mptr, d_mem, d_coef, d_out);
  EVOC_HIP(hipGetLastError());
}

void partition_means_device(const float* dX, int64_t ld, int64_t d, const int64_t* h_mptr,
                            const int32_t* h_mem, const float* h_coef, int64_t np, int64_t n,
                            float* h_out, hipStream_t s, DevPool& pool) {
  EVOC_CHECK(np >= 1 && n >= 1 && d >= 1, "partition_means: empty input");
  EVOC_CHECK(h_mptr[np] == n, "partition_means: member list does not cover the nodes");
  Buf<int64_t> mptr(pool, np + 1);
  Buf<int32_t> mem(pool, n);
  Buf<float> coef(pool, np), out(pool, np * d);
  EVOC_HIP(hipMemcpyAsync(mptr.p, h_mptr, (np + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
  EVOC_HIP(hipMemcpyAsync(mem.p, h_mem, n * sizeof(int32_t), hipMemcpyHostToDevice, s));
  EVOC_HIP(hipMemcpyAsync(coef.p, h_coef, np * sizeof(float), hipMemcpyHostToDevice, s));
  partition_means_dev(dX, ld, d, mptr.p, mem.p, coef.p, np, out.p, s);
  EVOC_HIP(hipMemcpyAsync(h_out, out.p, np * d * sizeof(float), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
}

}  // namespace evoc

// ---------------------------------------------------------------------------------
// Label votes and expansion of the input level (label_prop.rs:45-86, the expansion of
// label_prop_init_inner), one wavefront per node. Neighbour lists are walked in chunks
// of 64 slots; order-dependent sums are rebuilt from shuffles in CSR order, so every
// f32 value equals the sequential host formulation.
namespace evoc {
namespace {

constexpr int LP_WAVES = 4;  // wavefronts per workgroup

// One Jacobi step: an unlabelled node takes the label of largest summed edge weight
// (labels in first-appearance order, each label's weights added in CSR order, the
// running best starting at 1.0, ties to the later label), or stays unlabelled. A
// wavefront covers 64 consecutive nodes: labelled ones are copied lane-wise, and the
// unlabelled ones are taken one at a time by the whole wavefront.
__global__ __launch_bounds__(64 * LP_WAVES) void lp_step_kernel(
    const int64_t* __restrict__ indptr, const int32_t* __restrict__ col,
    const float* __restrict__ val, int64_t n, const int64_t* __restrict__ lab_in,
    int64_t* __restrict__ lab_out) {
  const int lane = threadIdx.x & 63;
  const int64_t i0 = ((int64_t)blockIdx.x * LP_WAVES + (threadIdx.x >> 6)) * 64;
  if (i0 >= n) return;  // whole wavefronts
  const int64_t me = i0 + lane;
  const int64_t lme = me < n ? lab_in[me] : 0;
  if (me < n && lme >= 0) lab_out[me] = lme;
  // the row bounds and first 64 slots (neighbour labels, weights) of the next unlabelled
  // node are loaded while the current one is voted
  int64_t ns0 = 0, ns1 = 0, nlab = -1;
  float nval = 0.0f;
  auto fetch = [&](int64_t j) {
    ns0 = indptr[j];
    ns1 = indptr[j + 1];
    const int64_t p = ns0 + lane;
    nlab = p < ns1 ? lab_in[col[p]] : -1;
    nval = p < ns1 ? val[p] : 0.0f;
  };
  unsigned long long todo = __ballot(me < n && lme < 0);
  if (todo) fetch(i0 + (__ffsll((long long)todo) - 1));
  for (; todo != 0; todo &= todo - 1) {
    const int64_t i = i0 + (__ffsll((long long)todo) - 1);
    const int64_t s0 = ns0, s1 = ns1, lab0 = nlab;
    const float val0 = nval;
    const unsigned long long rest = todo & (todo - 1);
    if (rest) fetch(i0 + (__ffsll((long long)rest) - 1));
    float top = 1.0f;
    int64_t best = -1;
    for (int64_t pb = s0; pb < s1; pb += 64) {
      const int64_t p = pb + lane;
      const bool pv = p < s1;
      const int64_t lp = pb == s0 ? lab0 : (pv ? lab_in[col[p]] : -1);
      if (__ballot(lp >= 0) == 0) continue;  // no label in this chunk: nothing to vote
      bool first = lp >= 0;
      float vw = pb == s0 ? val0 : (pv ? val[p] : 0.0f);
      for (int64_t qb = s0; qb < s1; qb += 64) {
        const int64_t q = qb + lane;
        const int64_t lq = qb == s0 ? lab0 : (q < s1 ? lab_in[col[q]] : -1);
        const float wq = qb == s0 ? val0 : (q < s1 ? val[q] : 0.0f);
        // only labelled slots can match a lane's label (an unlabelled lane's sums are
        // never read), so the walk visits those alone, in slot order
        for (unsigned long long lm = __ballot(lq >= 0); lm != 0; lm &= lm - 1) {
          const int t = __ffsll((long long)lm) - 1;
          const int64_t lt = __shfl(lq, t, 64);
          const float wt = __shfl(wq, t, 64);
          const int64_t qi = qb + t;
          if (lt == lp) {
            if (qi < p) first = false;
            else if (qi > p) vw = vw + wt;
          }
        }
      }
      // the chunk's first-appearance labels, in order: the sequential best update
      const unsigned long long fm = __ballot(first);
      for (unsigned long long m = fm; m; m &= m - 1) {
        const int t = __ffsll((long long)m) - 1;
        const float w = __shfl(vw, t, 64);
        const int64_t l = __shfl(lp, t, 64);
        if (w > top) {
          top = w;
          best = l;
        } else if (w == top) {
          best = l;
        }
      }
    }
    if (lane == 0) lab_out[i] = best;
  }
}

// ---- long rows (coarse graphs): one wavefront per node and an LDS sort of its slots
// Slots are keyed (label or partition << 16 | slot offset), so that after an ascending
// sort each label's slots are contiguous and in slot order: the per-label sums are the
// sequential adds of the CSR formulation, and the first slot of a run is the label's first
// appearance. Labels (seed indices) and partitions are below 16384, and rows up to
// LP_LONG_MAX slots fit the sort.
constexpr int LP_LONG_MAX = 16384;

// ascending bitonic sort of cnt (a power of two) keys with their values, one wavefront
// (the workgroup is one wavefront, so a barrier is only an ordering point)
__device__ void wave_sort(uint32_t* key, uint32_t* val, int cnt) {
  const int lane = threadIdx.x;
  for (int k = 2; k <= cnt; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = lane; t < cnt / 2; t += 64) {
        const int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
        const int l = i + j;
        const uint32_t a = key[i], b = key[l];
        if ((a > b) == ((i & k) == 0)) {
          key[i] = b;
          key[l] = a;
          const uint32_t va = val[i];
          val[i] = val[l];
          val[l] = va;
        }
      }
      __syncthreads();
    }
}

// the slots of row [s0, s1) with a key (16-bit id >= 0) into LDS, in slot order; returns
// their number, the rest of the power-of-two padding filled with the largest key
template <typename F>
__device__ int gather_keys(int64_t s0, int64_t s1, uint32_t* key, uint32_t* val, F&& slot) {
  const int lane = threadIdx.x;
  int m = 0;
  for (int64_t pb = s0; pb < s1; pb += 64) {
    const int64_t p = pb + lane;
    int32_t id = -1;
    float v = 0.0f;
    if (p < s1) slot(p, id, v);
    const unsigned long long b = __ballot(id >= 0);
    if (id >= 0) {
      const int pos = m + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
      key[pos] = ((uint32_t)id << 16) | (uint32_t)(p - s0);
      val[pos] = __float_as_uint(v);
    }
    m += __popcll(b);
  }
  int cnt = 1;
  while (cnt < m) cnt <<= 1;
  for (int e = m + lane; e < cnt; e += 64) {
    key[e] = 0xffffffffu;
    val[e] = 0u;
  }
  __syncthreads();
  wave_sort(key, val, cnt);
  return m;
}

// lp_step_kernel for rows of more than 64 slots: the label of largest summed weight (ties
// to the later first appearance) when that sum reaches the running best's start of 1.0.
__global__ __launch_bounds__(64) void lp_step_long_kernel(
    const int64_t* __restrict__ indptr, const int32_t* __restrict__ col,
    const float* __restrict__ val, int64_t n, const int64_t* __restrict__ lab_in,
    int64_t* __restrict__ lab_out) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lsm[];
  const int lane = threadIdx.x;
  const int64_t i = blockIdx.x;
  const int64_t li = lab_in[i];
  if (li >= 0) {
    if (lane == 0) lab_out[i] = li;
    return;
  }
  const int64_t s0 = indptr[i], s1 = indptr[i + 1];
  const int cap = (int)(s1 - s0) <= 1 ? 1 : 1 << (32 - __clz((int)(s1 - s0) - 1));
  uint32_t* key = lsm;
  uint32_t* vv = lsm + cap;
  const int m = gather_keys(s0, s1, key, vv, [&](int64_t p, int32_t& id, float& v) {
    const int64_t l = lab_in[col[p]];
    if (l >= 0) {
      id = (int32_t)l;
      v = val[p];
    }
  });
  float bw = -INFINITY;
  int bfirst = -1, blab = -1;
  for (int e = lane; e < m; e += 64) {
    const uint32_t ke = key[e];
    if (e > 0 && (key[e - 1] >> 16) == (ke >> 16)) continue;  // not the label's first slot
    float w = __uint_as_float(vv[e]);
    for (int f = e + 1; f < m && (key[f] >> 16) == (ke >> 16); ++f) w = w + __uint_as_float(vv[f]);
    const int first = (int)(ke & 0xffffu);
    if (w > bw || (w == bw && first > bfirst)) {
      bw = w;
      bfirst = first;
      blab = (int)(ke >> 16);
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const float ow = __shfl_xor(bw, o, 64);
    const int of = __shfl_xor(bfirst, o, 64), ol = __shfl_xor(blab, o, 64);
    if (ow > bw || (ow == bw && of > bfirst)) {
      bw = ow;
      bfirst = of;
      blab = ol;
    }
  }
  if (lane == 0) lab_out[i] = bw >= 1.0f ? (int64_t)blab : -1;
}

// exp_sums_kernel + exp_node_kernel for rows of more than 64 slots: the per-partition
// sums in slot order and the distinct partitions in ascending order come from the sort.
__global__ __launch_bounds__(64) void exp_long_kernel(
    const int64_t* __restrict__ indptr, int64_t n, int dim, const int32_t* __restrict__ part,
    const float* __restrict__ sa, const float* __restrict__ layout,
    const int32_t* __restrict__ ta, const float* __restrict__ tv, int32_t* __restrict__ ra,
    float* __restrict__ rv, float* __restrict__ res) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lsm[];
  const int lane = threadIdx.x;
  const int64_t i = blockIdx.x;
  const int64_t s0 = indptr[i], s1 = indptr[i + 1];
  const int cap = (int)(s1 - s0) <= 1 ? 1 : 1 << (32 - __clz((int)(s1 - s0) - 1));
  uint32_t* key = lsm;
  uint32_t* vv = lsm + cap;
  const int m = gather_keys(s0, s1, key, vv, [&](int64_t p, int32_t& id, float& v) {
    id = ta[p];
    v = tv[p];
  });
  int ndist = 0;
  for (int eb = 0; eb < m; eb += 64) {
    const int e = eb + lane;
    const bool head = e < m && (e == 0 || (key[e - 1] >> 16) != (key[e] >> 16));
    const unsigned long long b = __ballot(head);
    if (head) {
      float acc = __uint_as_float(vv[e]);
      for (int f = e + 1; f < m && (key[f] >> 16) == (key[e] >> 16); ++f)
        acc = acc + __uint_as_float(vv[f]);
      const int r = ndist + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
      ra[s0 + r] = (int32_t)(key[e] >> 16);
      rv[s0 + r] = acc;
    }
    ndist += __popcll(b);
  }
  __threadfence_block();  // ra/rv written by other lanes of this wavefront
  __syncthreads();
  float l1 = 0.0f;
  for (int r = 0; r < ndist; ++r) l1 = l1 + fabsf(rv[s0 + r]);
  const float inv = l1 > 0.0f ? 1.0f / l1 : 1.0f;
  if (lane < dim) {
    float ex = 0.0f;
    for (int r = 0; r < ndist; ++r) {
      const float v0 = rv[s0 + r];
      const float v = l1 > 0.0f ? v0 * inv : v0;
      const float t = v * layout[(int64_t)ra[s0 + r] * dim + lane];
      ex = ex + t;
    }
    const int32_t pi = part[i];
    const float sv = sa[pi];
    const float dv = sv > 0.0f ? sv * (1.0f / fabsf(sv)) : sv;
    const float t = dv * layout[(int64_t)pi * dim + lane];
    const float dr = 0.0f + t;
    res[i * dim + lane] = (ex + dr) * 0.5f;
  }
}

__global__ void row_length_max_kernel(const int64_t* __restrict__ indptr, int64_t n,
                                      unsigned long long* __restrict__ mx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = i < n ? (unsigned long long)(indptr[i + 1] - indptr[i]) : 0ull;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const unsigned long long w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  if ((threadIdx.x & 63) == 0 && v) atomicMax(mx, v);
}
__global__ void zero_u64_kernel(unsigned long long* p) { *p = 0ull; }

// Expansion, pass 1 per node: for every slot the partition a of the neighbour and the
// term sym(i, j) * sa[a] when G(j, i) exists (a = -1 otherwise).
__global__ __launch_bounds__(256) void exp_terms_kernel(
    const int64_t* __restrict__ indptr, const int32_t* __restrict__ col,
    const float* __restrict__ val, int64_t n, const int32_t* __restrict__ part,
    const float* __restrict__ sa, int32_t* __restrict__ ta, float* __restrict__ tv) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * LP_WAVES + (threadIdx.x >> 6);
  if (i >= n) return;
  const int64_t s0 = indptr[i], s1 = indptr[i + 1];
  for (int64_t p = s0 + lane; p < s1; p += 64) {
    const int32_t j = col[p];
    // G^T(i, j) = G(j, i): binary search in row j (columns ascending)
    int64_t lo = indptr[j], hi = indptr[j + 1];
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (col[mid] < (int32_t)i) lo = mid + 1;
      else hi = mid;
    }
    int32_t a = -1;
    float term = 0.0f;
    if (lo < indptr[j + 1] && col[lo] == (int32_t)i) {
      const float sym = val[p] * val[lo];
      a = part[j];
      term = sym * sa[a];
    }
    ta[p] = a;
    tv[p] = term;
  }
}

// Expansion, pass 2 per node: per-partition sums in CSR order, carried by the first
// slot of each partition (flag tf), sums in tacc.
__global__ __launch_bounds__(256) void exp_sums_kernel(const int64_t* __restrict__ indptr,
                                                       int64_t n, const int32_t* __restrict__ ta,
                                                       const float* __restrict__ tv,
                                                       int32_t* __restrict__ tf,
                                                       float* __restrict__ tacc) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * LP_WAVES + (threadIdx.x >> 6);
  if (i >= n) return;
  const int64_t s0 = indptr[i], s1 = indptr[i + 1];
  for (int64_t pb = s0; pb < s1; pb += 64) {
    const int64_t p = pb + lane;
    const int32_t ap = p < s1 ? ta[p] : -1;
    bool first = ap >= 0;
    float acc = p < s1 ? tv[p] : 0.0f;
    for (int64_t qb = s0; qb < s1; qb += 64) {
      const int64_t q = qb + lane;
      const int32_t aq = q < s1 ? ta[q] : -1;
      const float vq = q < s1 ? tv[q] : 0.0f;
      const int tn = (int)min<int64_t>(64, s1 - qb);
      for (int t = 0; t < tn; ++t) {
        const int32_t at = __shfl(aq, t, 64);
        const float vt = __shfl(vq, t, 64);
        const int64_t qi = qb + t;
        if (at == ap) {
          if (qi < p) first = false;
          else if (qi > p) acc = acc + vt;
        }
      }
    }
    if (p < s1) {
      tf[p] = first ? 1 : 0;
      tacc[p] = acc;
    }
  }
}

// Expansion, pass 3 per node: the distinct partitions in ascending order (rank = number
// of distinct smaller ids) into ra/rv, then the L1 norm and the layout product summed in
// that order, averaged with the layout of the node's own partition.
__global__ __launch_bounds__(256) void exp_node_kernel(
    const int64_t* __restrict__ indptr, int64_t n, int dim, const int32_t* __restrict__ part,
    const float* __restrict__ sa, const float* __restrict__ layout,
    const int32_t* __restrict__ ta, const int32_t* __restrict__ tf,
    const float* __restrict__ tacc, int32_t* __restrict__ ra, float* __restrict__ rv,
    float* __restrict__ res) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * LP_WAVES + (threadIdx.x >> 6);
  if (i >= n) return;
  const int64_t s0 = indptr[i], s1 = indptr[i + 1];
  int ndist = 0;
  for (int64_t pb = s0; pb < s1; pb += 64) {
    const int64_t p = pb + lane;
    const bool first = p < s1 && tf[p] != 0;
    const int32_t ap = p < s1 ? ta[p] : -1;
    int rank = 0;
    for (int64_t qb = s0; qb < s1; qb += 64) {
      const int64_t q = qb + lane;
      const int32_t aq = (q < s1 && tf[q] != 0) ? ta[q] : INT32_MAX;
      const int tn = (int)min<int64_t>(64, s1 - qb);
      for (int t = 0; t < tn; ++t) rank += (__shfl(aq, t, 64) < ap) ? 1 : 0;
    }
    if (first) {
      ra[s0 + rank] = ap;
      rv[s0 + rank] = tacc[p];
    }
    ndist += __popcll(__ballot(first));
  }
  __threadfence_block();  // ra/rv written by other lanes of this wavefront
  float l1 = 0.0f;
  for (int r = 0; r < ndist; ++r) l1 = l1 + fabsf(rv[s0 + r]);
  const float inv = l1 > 0.0f ? 1.0f / l1 : 1.0f;
  if (lane < dim) {
    float ex = 0.0f;
    for (int r = 0; r < ndist; ++r) {
      const float v0 = rv[s0 + r];
      const float v = l1 > 0.0f ? v0 * inv : v0;
      const float t = v * layout[(int64_t)ra[s0 + r] * dim + lane];
      ex = ex + t;
    }
    const int32_t pi = part[i];
    const float sv = sa[pi];
    const float dv = sv > 0.0f ? sv * (1.0f / fabsf(sv)) : sv;
    const float t = dv * layout[(int64_t)pi * dim + lane];
    const float dr = 0.0f + t;
    res[i * dim + lane] = (ex + dr) * 0.5f;
  }
}

// Column means of the n x dim result, each column summed sequentially in row order
// (the reference's iterator sum). The workgroup streams 64 KB chunks of rows into LDS,
// transposed, double-buffered; one lane per column then runs the dependent adds out of
// LDS while the next chunk's loads are in flight.
constexpr int CS_THREADS = 512;
constexpr int CS_LOADERS = CS_THREADS - 64;  // wavefront 0 sums, the others stage chunks
constexpr int CS_FLOATS = 16384;  // floats per chunk buffer (64 KB)

// LDS pitch of one column: 4 mod 64 words, so that the summing lanes' 16-byte reads start
// in distinct groups of four banks; rows per chunk = pitch - 4 (a multiple of 64)
__host__ __device__ inline int cs_pitch(int dim) { return (CS_FLOATS / dim - 4) / 64 * 64 + 4; }
__host__ __device__ inline int cs_rows(int dim) { return cs_pitch(dim) - 4; }

__global__ __launch_bounds__(CS_THREADS) void exp_colsum_kernel(const float* __restrict__ res,
                                                                int64_t n, int dim, float inv_n,
                                                                float* __restrict__ mean) {
  extern __shared__ __attribute__((aligned(16))) float cs_buf[];  // [2][dim][pitch]
  const int rows = cs_rows(dim);
  const int pitch = cs_pitch(dim);
  const int t = threadIdx.x;
  const int64_t nchunks = (n + rows - 1) / rows;
  // wavefront 0 only sums, so that its dependent adds never wait on its own share of the
  // staging; the other wavefronts load the next chunk and store it transposed
  constexpr int PER = (CS_FLOATS + CS_LOADERS - 1) / CS_LOADERS;  // floats per loader
  const bool loader = t >= 64;
  const int lt = t - 64;
  float stage[PER];
  auto load = [&](int64_t k) {
    if (!loader) return;
    const int64_t base = k * rows * dim;
    const int64_t lim = n * dim - base;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int q = u * CS_LOADERS + lt;
      stage[u] = (q < rows * dim && q < lim) ? res[base + q] : 0.0f;
    }
  };
  auto store = [&](int buf) {
    if (!loader) return;
    float* B = cs_buf + buf * dim * pitch;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int q = u * CS_LOADERS + lt;
      if (q < rows * dim) B[(q % dim) * pitch + q / dim] = stage[u];
    }
  };
  float sum = 0.0f;
  load(0);
  store(0);
  for (int64_t k = 0; k < nchunks; ++k) {
    __syncthreads();
    if (k + 1 < nchunks) load(k + 1);
    if (t < dim) {
      const float4* B = reinterpret_cast<const float4*>(cs_buf + (k & 1) * dim * pitch + t * pitch);
      const int cnt = (int)min<int64_t>(rows, n - k * rows);
      int r = 0;
      for (; r + 16 <= cnt; r += 16) {  // four 16-byte reads ahead of the dependent adds
        const float4 a = B[r / 4], b = B[r / 4 + 1], c = B[r / 4 + 2], d = B[r / 4 + 3];
        sum = sum + a.x; sum = sum + a.y; sum = sum + a.z; sum = sum + a.w;
        sum = sum + b.x; sum = sum + b.y; sum = sum + b.z; sum = sum + b.w;
        sum = sum + c.x; sum = sum + c.y; sum = sum + c.z; sum = sum + c.w;
        sum = sum + d.x; sum = sum + d.y; sum = sum + d.z; sum = sum + d.w;
      }
      const float* Bs = reinterpret_cast<const float*>(B);
      for (; r < cnt; ++r) sum = sum + Bs[r];
    }
    if (k + 1 < nchunks) store((k + 1) & 1);
  }
  if (t < dim) mean[t] = sum * inv_n;
}
__global__ void exp_center_kernel(float* __restrict__ res, int64_t n, int dim,
                                  const float* __restrict__ mean, float* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * dim) return;
  const float x = res[q] - mean[q % dim];
  out[q] = 0.1f * x;
}

}  // namespace

// the longest row of a device CSR
static int64_t max_row_length(const int64_t* d_ptr, int64_t n, hipStream_t s, DevPool& pool) {
  if (n <= 0) return 0;
  Buf<unsigned long long> mx(pool, 1);
  zero_u64_kernel<<<1, 1, 0, s>>>(mx.p);
  row_length_max_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(d_ptr, n, mx.p);
  EVOC_HIP(hipGetLastError());
  unsigned long long h = 0;
  EVOC_HIP(hipMemcpyAsync(&h, mx.p, sizeof(h), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
  return (int64_t)h;
}
// the LDS of the long-row kernels for rows of up to maxdeg slots (keys and values)
static size_t long_row_lds(int64_t maxdeg) {
  int64_t cap = 1;
  while (cap < maxdeg) cap <<= 1;
  return (size_t)cap * 8;
}

void expansion_device(const int64_t* d_indptr, const int32_t* d_col, const float* d_val,
                      int64_t n, int64_t nnz, int dim, const int32_t* d_part, const float* d_sa,
                      int64_t np, const float* h_layout, float* d_out, hipStream_t s,
                      DevPool& pool) {
  const int64_t m = std::max<int64_t>(nnz, 1);
  Buf<int32_t> ta(pool, m), tf(pool, m), ra(pool, m);
  Buf<float> layout(pool, np * dim), tv(pool, m), tacc(pool, m), rv(pool, m), res(pool, n * dim),
      mean(pool, dim);
  EVOC_HIP(hipMemcpyAsync(layout.p, h_layout, np * dim * sizeof(float), hipMemcpyHostToDevice, s));
  const unsigned g = (unsigned)ceil_div(n, LP_WAVES);
  exp_terms_kernel<<<g, 64 * LP_WAVES, 0, s>>>(d_indptr, d_col, d_val, n, d_part, d_sa, ta.p,
                                               tv.p);
  // rows of more than 64 slots on graphs of at most 16384 nodes (the coarse levels): a
  // wavefront per node sorts its slots
  const int64_t maxdeg = max_row_length(d_indptr, n, s, pool);
  if (maxdeg > 64 && n <= LP_LONG_MAX && np <= LP_LONG_MAX) {
    const size_t lds = long_row_lds(maxdeg);
    EVOC_HIP(hipFuncSetAttribute((const void*)exp_long_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    exp_long_kernel<<<(unsigned)n, 64, lds, s>>>(d_indptr, n, dim, d_part, d_sa, layout.p, ta.p,
                                                tv.p, ra.p, rv.p, res.p);
  } else {
    exp_sums_kernel<<<g, 64 * LP_WAVES, 0, s>>>(d_indptr, n, ta.p, tv.p, tf.p, tacc.p);
    exp_node_kernel<<<g, 64 * LP_WAVES, 0, s>>>(d_indptr, n, dim, d_part, d_sa, layout.p, ta.p,
                                                tf.p, tacc.p, ra.p, rv.p, res.p);
  }
  {
    const size_t lds = 2 * (size_t)dim * cs_pitch(dim) * sizeof(float);
    EVOC_HIP(hipFuncSetAttribute((const void*)exp_colsum_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    exp_colsum_kernel<<<1, CS_THREADS, lds, s>>>(res.p, n, dim, (float)(1.0 / (double)n), mean.p);
  }
  exp_center_kernel<<<(unsigned)ceil_div(n * dim, 256), 256, 0, s>>>(res.p, n, dim, mean.p, d_out);
  EVOC_HIP(hipGetLastError());
  EVOC_HIP(hipStreamSynchronize(s));
}

}  // namespace evoc

// ---------------------------------------------------------------------------------
// Partition and coarse graph of the input level in device memory: label_prop_loop
// (label_prop.rs:202-227: seeds, Jacobi votes, label_outliers, remap_labels), the
// member lists and column norms of the partition map, and R^T_norm G R (the
// per-partition accumulation of coarsen, columns ascending) built from two stable
// sorts, with every order-dependent f32 sum walked in the host formulation's order.
namespace evoc {
namespace {

__global__ void lp_fill_kernel(int64_t n, int64_t* __restrict__ lab) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) lab[i] = -1;
}
__global__ void lp_seed_kernel(int64_t ns, const int64_t* __restrict__ pos,
                               const int64_t* __restrict__ t, int64_t* __restrict__ lab) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < ns) lab[pos[q]] = t[q];
}
// flag[i] = 1 for unlabelled nodes; flag[n] = 0 so that the scan's last entry is the count
__global__ void lp_unlabelled_kernel(int64_t n, const int64_t* __restrict__ lab,
                                     int64_t* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) flag[i] = (i < n && lab[i] < 0) ? 1 : 0;
}
__global__ void lp_compact_kernel(int64_t n, const int64_t* __restrict__ flag,
                                  const int64_t* __restrict__ rank, int64_t* __restrict__ list) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i]) list[rank[i]] = i;
}

// label_outliers (label_prop.rs:105-140) runs the outliers in node order, so when
// outlier i searches, exactly the outliers below i already carry labels. Each thread
// runs one outlier's depth-first search under that rule and records a label, or a
// reference to the earlier outlier whose final label it takes, or (search exhausted)
// its random label. Only the top 64 stack entries can ever be popped within the 64-pop
// limit, so a 64-entry ring holds the stack exactly.
__global__ void lp_outlier_walk_kernel(const int64_t* __restrict__ ptr,
                                       const int32_t* __restrict__ col,
                                       const int64_t* __restrict__ lab,
                                       const int64_t* __restrict__ list, int64_t m,
                                       uint64_t seed, int64_t maxl, int32_t* __restrict__ kind,
                                       int64_t* __restrict__ val) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int64_t i = list[k];
  int64_t stk[64];
  int top = 0, size = 0;
  stk[0] = i;
  top = 1;
  size = 1;
  int kd = 0;
  int64_t v = 0;
  for (int it = 0; it < 64 && size > 0 && kd == 0; ++it) {
    --top;
    --size;
    const int64_t node = stk[top & 63];
    for (int64_t p = ptr[node]; p < ptr[node + 1]; ++p) {
      const int64_t j = col[p];
      const int64_t lj = lab[j];
      if (lj >= 0) {
        kd = 1;
        v = lj;
        break;
      }
      if (j < i) {  // an earlier outlier: labelled by now
        kd = 2;
        v = j;
        break;
      }
      stk[top & 63] = j;
      ++top;
      ++size;
    }
  }
  if (kd == 0) {
    const uint32_t r = Philox::first(seed, (uint32_t)i, (uint32_t)(i >> 32), 0, TAG_LPOL);
    v = (int64_t)Philox::range(r, (uint64_t)maxl + 1);
    kd = 1;
  }
  kind[k] = kd;
  val[k] = v;
}
// references point to strictly smaller outliers, so every chain ends in a label
__global__ void lp_outlier_resolve_kernel(const int64_t* __restrict__ list, int64_t m,
                                          const int64_t* __restrict__ rank,
                                          const int32_t* __restrict__ kind,
                                          const int64_t* __restrict__ val,
                                          int64_t* __restrict__ lab) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  int64_t q = k;
  while (kind[q] == 2) q = rank[val[q]];
  lab[list[k]] = val[q];
}

// remap_labels: the labels present, in ascending order, become 0..np-1
__global__ void lp_present_kernel(int64_t n, const int64_t* __restrict__ lab,
                                  int64_t* __restrict__ present) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && lab[i] >= 0) present[lab[i]] = 1;
}
__global__ void lp_part_kernel(int64_t n, const int64_t* __restrict__ lab,
                               const int64_t* __restrict__ rank, int32_t* __restrict__ part,
                               uint64_t* __restrict__ key, uint32_t* __restrict__ idx,
                               int64_t* __restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t a = lab[i] >= 0 ? (int32_t)rank[lab[i]] : 0;  // every node is labelled here
  part[i] = a;
  key[i] = (uint64_t)a;
  idx[i] = (uint32_t)i;
  atomicAdd(reinterpret_cast<unsigned long long*>(&count[a]), 1ull);
}
__global__ void lp_idx_to_mem_kernel(int64_t n, const uint32_t* __restrict__ idx,
                                     int32_t* __restrict__ mem) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) mem[i] = (int32_t)idx[i];
}
// x added cnt times to 0.0f in f32, one rounding per add, for a positive normal x whose
// partial sums stay finite. Within one binade [2^(e-1), 2^e) of the running sum every add
// moves it by the same multiple of the binade's ulp u (x/u rounded to nearest; on a tie the
// first add makes the sum's last bit even and later adds keep it so), so whole runs of adds
// are taken in one double-precision step and only the add that leaves a binade is done in f32.
__device__ float repeated_sum(float x, int64_t cnt) {
  if (cnt <= 0) return 0.0f;
  float s = x;  // 0 + x is exact
  int64_t k = 1;
  while (k < cnt) {
    int e;
    frexp((double)s, &e);  // s in [2^(e-1), 2^e)
    const double u = ldexp(1.0, e - 24), top = ldexp(1.0, e);
    const double q = (double)x / u, f = floor(q);
    double step;
    if (q - f == 0.5) {
      s = s + x;
      ++k;
      int e2;
      frexp((double)s, &e2);
      if (e2 != e || k >= cnt) continue;
      step = (fmod(f, 2.0) == 0.0 ? f : f + 1.0) * u;
    } else {
      step = (q - f < 0.5 ? f : f + 1.0) * u;
    }
    if (step <= 0.0) break;  // x below half an ulp: the sum no longer moves
    // adds j = 1..m start below top - x, so each lands on s + j*step (at most top)
    const double room = top - (double)x - (double)s;
    int64_t m = room > 0.0 ? (int64_t)ceil(room / step) : 0;
    while (m > 0 && (double)(m - 1) * step >= room) --m;
    while ((double)m * step < room) ++m;
    if (m > cnt - k) m = cnt - k;
    s = (float)((double)s + (double)m * step);
    k += m;
    if (k < cnt) {
      s = s + x;
      ++k;
    }
  }
  return s;
}

// column norm 1/sqrt(|a|) of the partition map (1.0f added |a| times, as the host sums
// it) and the L1 row normalisation of R^T_norm (|sa| added |a| times in order)
__global__ void lp_norms_kernel(int64_t np, const int64_t* __restrict__ mptr,
                                float* __restrict__ sa, float* __restrict__ coef) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= np) return;
  const int64_t cnt = mptr[a + 1] - mptr[a];
  const float sq = (float)min<int64_t>(cnt, 16777216);  // f32 count saturates at 2^24
  const float nrm = sqrtf(sq);
  const float s = nrm > 0.0f ? 1.0f / nrm : 1.0f;
  const float l1 = repeated_sum(fabsf(s), cnt);  // s is in [2^-12, 1]
  sa[a] = s;
  coef[a] = l1 > 0.0f ? s * (1.0f / l1) : s;
}

// coarsening, step 1: entry p of row i gets key (part[i], col[p]); the stable sort
// keeps CSR order (= member order, then row order) within a key
__global__ void cg_key1_kernel(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                               int64_t n, const int32_t* __restrict__ part,
                               uint64_t* __restrict__ key, uint32_t* __restrict__ idx) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  const uint64_t a = (uint64_t)part[i] * (uint64_t)n;
  for (int64_t p = ptr[i] + lane; p < ptr[i + 1]; p += 64) {
    key[p] = a + (uint64_t)col[p];
    idx[p] = (uint32_t)p;
  }
}
// flag[q] = 1 where a run of equal keys starts; flag[m] = 0
__global__ void cg_heads_kernel(int64_t m, const uint64_t* __restrict__ key,
                                int64_t* __restrict__ flag) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q <= m) flag[q] = (q < m && (q == 0 || key[q] != key[q - 1])) ? 1 : 0;
}
__global__ void cg_starts_kernel(int64_t m, const int64_t* __restrict__ flag,
                                 const int64_t* __restrict__ rank, int64_t* __restrict__ start) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < m && flag[q]) start[rank[q]] = q;
}
// step 2: per (a, j) run, the sum of sa[a] * G(i, j) in run order; the run then gets key
// (a, part[j]) for the second stable sort (j stays ascending within it)
__global__ void cg_run_sum_kernel(int64_t nrun, int64_t m, int64_t n, int64_t np,
                                  const int64_t* __restrict__ start,
                                  const uint64_t* __restrict__ key, const uint32_t* __restrict__ idx,
                                  const float* __restrict__ val, const float* __restrict__ sa,
                                  const int32_t* __restrict__ part, float* __restrict__ acc,
                                  uint64_t* __restrict__ key2, uint32_t* __restrict__ idx2) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrun) return;
  const int64_t q0 = start[r], q1 = r + 1 < nrun ? start[r + 1] : m;
  const uint64_t k = key[q0];
  const int64_t a = (int64_t)(k / (uint64_t)n), j = (int64_t)(k % (uint64_t)n);
  const float s = sa[a];
  float x = s * val[idx[q0]];
  for (int64_t q = q0 + 1; q < q1; ++q) {
    const float pr = s * val[idx[q]];
    x = x + pr;
  }
  acc[r] = x;
  key2[r] = (uint64_t)a * (uint64_t)np + (uint64_t)part[j];
  idx2[r] = (uint32_t)r;
}
// the (a, j) sums in the order of the second sort, so that runs are read contiguously
__global__ void cg_gather_kernel(int64_t nrun, const uint32_t* __restrict__ idx2,
                                 const float* __restrict__ acc, float* __restrict__ accs) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nrun) accs[r] = acc[idx2[r]];
}
// step 3: per (a, b) run, the sum of the (a, j) sums in ascending j, clipped to [0, 1]
__global__ void cg_entry_kernel(int64_t nent, int64_t nrun, int64_t np,
                                const int64_t* __restrict__ start,
                                const uint64_t* __restrict__ key2, const float* __restrict__ accs,
                                int32_t* __restrict__ ccol, float* __restrict__ cval,
                                int64_t* __restrict__ count) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nent) return;
  const int64_t r0 = start[e], r1 = e + 1 < nent ? start[e + 1] : nrun;
  const uint64_t k = key2[r0];
  float v = accs[r0] * 1.0f;
  int64_t r = r0 + 1;
  for (; r + 8 <= r1; r += 8) {  // loads ahead of the dependent adds
    float x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = accs[r + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const float pr = x[u] * 1.0f;
      v = v + pr;
    }
  }
  for (; r < r1; ++r) {
    const float pr = accs[r] * 1.0f;
    v = v + pr;
  }
  v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
  ccol[e] = (int32_t)(k % (uint64_t)np);
  cval[e] = v;
  atomicAdd(reinterpret_cast<unsigned long long*>(&count[k / (uint64_t)np]), 1ull);
}

// Embedding slots of a coarse graph (one per endpoint of every entry (i, j) with i < j,
// as the host builds them in row order): node x lists first the rows i < x holding (i,
// x), ascending, then its own entries j > x, ascending. Each slot gets the key
// (2 x + group) * np + other, so one sort by key lays the lists out in that order.
__global__ __launch_bounds__(256) void slot_key_kernel(const int64_t* __restrict__ ptr,
                                                       const int32_t* __restrict__ col,
                                                       int64_t np, uint64_t* __restrict__ key,
                                                       uint32_t* __restrict__ idx,
                                                       int32_t* __restrict__ erow,
                                                       int64_t* __restrict__ deg) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * LP_WAVES + (threadIdx.x >> 6);
  if (i >= np) return;
  const uint64_t none = (uint64_t)(2 * np) * (uint64_t)np;
  int64_t own = 0;
  for (int64_t p = ptr[i] + lane; p < ptr[i + 1]; p += 64) {
    const int32_t j = col[p];
    erow[p] = (int32_t)i;
    const bool up = (int64_t)j > i;
    key[2 * p] = up ? (uint64_t)(2 * i + 1) * (uint64_t)np + (uint64_t)j : none;
    key[2 * p + 1] = up ? (uint64_t)(2 * (int64_t)j) * (uint64_t)np + (uint64_t)i : none;
    idx[2 * p] = (uint32_t)(2 * p);
    idx[2 * p + 1] = (uint32_t)(2 * p + 1);
    if (up) {
      ++own;
      atomicAdd(reinterpret_cast<unsigned long long*>(&deg[j]), 1ull);
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) own += __shfl_xor(own, o, 64);
  if (lane == 0) atomicAdd(reinterpret_cast<unsigned long long*>(&deg[i]), (unsigned long long)own);
}
// slot s of the sorted order: the other endpoint and the entry's weight
__global__ void slot_gather_kernel(int64_t m, const uint32_t* __restrict__ idx,
                                   const int32_t* __restrict__ col, const int32_t* __restrict__ erow,
                                   const float* __restrict__ val, int32_t* __restrict__ other,
                                   float* __restrict__ w) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const uint32_t t = idx[q];
  const int64_t p = t >> 1;
  other[q] = (t & 1u) ? erow[p] : col[p];
  w[q] = val[p];
}

int bit_length(uint64_t x) {
  int b = 0;
  while (x) {
    ++b;
    x >>= 1;
  }
  return std::max(b, 1);
}

}  // namespace

void lp_top_partition(const int64_t* d_ptr, const int32_t* d_col, const float* d_val, int64_t n,
                      int64_t approx, int n_iter, uint64_t seed, LpTop& P, hipStream_t s,
                      DevPool& pool) {
  EVOC_CHECK(approx >= 1 && approx <= n, "label propagation: seed count out of range");
  EVOC_CHECK(n < ((int64_t)1 << 31), "label propagation: n must be below 2^31");
  // seeds: seed t lands on a random node; a later seed on the same node wins
  std::map<int64_t, int64_t> seeds;
  for (int64_t t = 0; t < approx; ++t) {
    const uint32_t r = Philox::first(seed, (uint32_t)t, (uint32_t)(t >> 32), 0, TAG_LPSD);
    seeds[(int64_t)Philox::range(r, (uint64_t)n)] = t;
  }
  std::vector<int64_t> hp, ht;
  int64_t maxl = -1;
  for (auto& kv : seeds) {
    hp.push_back(kv.first);
    ht.push_back(kv.second);
    maxl = std::max(maxl, kv.second);
  }
  const int64_t ns = (int64_t)hp.size();
  EVOC_CHECK(maxl >= 0 && maxl < approx, "label propagation: no seeds");
  Buf<int64_t> dpos(pool, ns), dt(pool, ns), la(pool, n), lb(pool, n);
  EVOC_HIP(hipMemcpyAsync(dpos.p, hp.data(), ns * sizeof(int64_t), hipMemcpyHostToDevice, s));
  EVOC_HIP(hipMemcpyAsync(dt.p, ht.data(), ns * sizeof(int64_t), hipMemcpyHostToDevice, s));
  const unsigned gn = (unsigned)ceil_div(n, 256);
  lp_fill_kernel<<<gn, 256, 0, s>>>(n, la.p);
  lp_seed_kernel<<<(unsigned)ceil_div(ns, 256), 256, 0, s>>>(ns, dpos.p, dt.p, la.p);
  // Jacobi votes (labelled nodes keep their label, so every seed label survives and
  // maxl is the largest seed label)
  const unsigned gw = (unsigned)ceil_div(n, 64 * LP_WAVES);
  // rows of more than 64 slots on graphs of at most 16384 nodes (the coarse levels; a row
  // is then at most 16384 slots and seed labels are below approx <= 16384): a wavefront
  // per node sorts its labelled slots
  const int64_t maxdeg = max_row_length(d_ptr, n, s, pool);
  const bool long_rows = maxdeg > 64 && n <= LP_LONG_MAX && approx <= LP_LONG_MAX;
  const size_t lds = long_rows ? long_row_lds(maxdeg) : 0;
  if (long_rows)
    EVOC_HIP(hipFuncSetAttribute((const void*)lp_step_long_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  for (int it = 0; it < n_iter; ++it) {
    if (long_rows)
      lp_step_long_kernel<<<(unsigned)n, 64, lds, s>>>(d_ptr, d_col, d_val, n, la.p, lb.p);
    else
      lp_step_kernel<<<gw, 64 * LP_WAVES, 0, s>>>(d_ptr, d_col, d_val, n, la.p, lb.p);
    std::swap(la.p, lb.p);
  }
  // outliers
  Buf<int64_t> flag(pool, n + 1), rank(pool, n + 1);
  lp_unlabelled_kernel<<<(unsigned)ceil_div(n + 1, 256), 256, 0, s>>>(n, la.p, flag.p);
  exclusive_scan_i64(flag.p, rank.p, n + 1, s, pool);
  int64_t m = 0;
  EVOC_HIP(hipMemcpyAsync(&m, rank.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
  if (m > 0) {
    Buf<int64_t> list(pool, m), oval(pool, m);
    Buf<int32_t> okind(pool, m);
    lp_compact_kernel<<<gn, 256, 0, s>>>(n, flag.p, rank.p, list.p);
    const unsigned gm = (unsigned)ceil_div(m, 64);
    lp_outlier_walk_kernel<<<gm, 64, 0, s>>>(d_ptr, d_col, la.p, list.p, m, seed + 2000, maxl,
                                             okind.p, oval.p);
    lp_outlier_resolve_kernel<<<gm, 64, 0, s>>>(list.p, m, rank.p, okind.p, oval.p, la.p);
  }
  // remap: labels are seed ids below approx
  Buf<int64_t> present(pool, approx + 1), lrank(pool, approx + 1);
  EVOC_HIP(hipMemsetAsync(present.p, 0, (approx + 1) * sizeof(int64_t), s));
  lp_present_kernel<<<gn, 256, 0, s>>>(n, la.p, present.p);
  exclusive_scan_i64(present.p, lrank.p, approx + 1, s, pool);
  EVOC_HIP(hipMemcpyAsync(&P.np, lrank.p + approx, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
  const int64_t np = P.np;
  // members of each partition in ascending node order: stable sort by partition
  P.part.alloc(pool, n);
  P.mem.alloc(pool, n);
  P.mptr.alloc(pool, np + 1);
  P.sa.alloc(pool, np);
  P.coef.alloc(pool, np);
  Buf<uint64_t> key(pool, n);
  Buf<uint32_t> idx(pool, n);
  Buf<int64_t> cnt(pool, np + 1);
  EVOC_HIP(hipMemsetAsync(cnt.p, 0, (np + 1) * sizeof(int64_t), s));
  lp_part_kernel<<<gn, 256, 0, s>>>(n, la.p, lrank.p, P.part.p, key.p, idx.p, cnt.p);
  radix_sort_pairs(key.p, idx.p, n, bit_length((uint64_t)(np - 1)), s, pool);
  lp_idx_to_mem_kernel<<<gn, 256, 0, s>>>(n, idx.p, P.mem.p);
  exclusive_scan_i64(cnt.p, P.mptr.p, np + 1, s, pool);
  lp_norms_kernel<<<(unsigned)ceil_div(np, 256), 256, 0, s>>>(np, P.mptr.p, P.sa.p, P.coef.p);
  EVOC_HIP(hipGetLastError());
}

void lp_top_coarsen(const int64_t* d_ptr, const int32_t* d_col, const float* d_val, int64_t n,
                    int64_t nnz, const LpTop& P, std::vector<int64_t>& cptr,
                    std::vector<int32_t>& ccol, std::vector<float>& cval, hipStream_t s,
                    DevPool& pool, Graph* keep, bool host_copy) {
  const int64_t np = P.np;
  cptr.assign(np + 1, 0);
  ccol.clear();
  cval.clear();
  if (nnz == 0) return;
  EVOC_CHECK(nnz < ((int64_t)1 << 32), "coarsening: more than 2^32 graph entries");
  Buf<uint64_t> key(pool, nnz);
  Buf<uint32_t> idx(pool, nnz);
  cg_key1_kernel<<<(unsigned)ceil_div(n, 4), 256, 0, s>>>(d_ptr, d_col, n, P.part.p, key.p, idx.p);
  radix_sort_pairs(key.p, idx.p, nnz, bit_length((uint64_t)np * (uint64_t)n - 1), s, pool);
  Buf<int64_t> flag(pool, nnz + 1), rank(pool, nnz + 1);
  cg_heads_kernel<<<(unsigned)ceil_div(nnz + 1, 256), 256, 0, s>>>(nnz, key.p, flag.p);
  exclusive_scan_i64(flag.p, rank.p, nnz + 1, s, pool);
  int64_t nrun = 0;
  EVOC_HIP(hipMemcpyAsync(&nrun, rank.p + nnz, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
  Buf<int64_t> start(pool, nrun);
  cg_starts_kernel<<<(unsigned)ceil_div(nnz, 256), 256, 0, s>>>(nnz, flag.p, rank.p, start.p);
  Buf<float> acc(pool, nrun);
  Buf<uint64_t> key2(pool, nrun);
  Buf<uint32_t> idx2(pool, nrun);
  cg_run_sum_kernel<<<(unsigned)ceil_div(nrun, 256), 256, 0, s>>>(
      nrun, nnz, n, np, start.p, key.p, idx.p, d_val, P.sa.p, P.part.p, acc.p, key2.p, idx2.p);
  radix_sort_pairs(key2.p, idx2.p, nrun, bit_length((uint64_t)np * (uint64_t)np - 1), s, pool);
  cg_heads_kernel<<<(unsigned)ceil_div(nrun + 1, 256), 256, 0, s>>>(nrun, key2.p, flag.p);
  exclusive_scan_i64(flag.p, rank.p, nrun + 1, s, pool);
  int64_t nent = 0;
  EVOC_HIP(hipMemcpyAsync(&nent, rank.p + nrun, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
  Buf<int64_t> start2(pool, nent), cnt(pool, np + 1), dptr(pool, np + 1);
  cg_starts_kernel<<<(unsigned)ceil_div(nrun, 256), 256, 0, s>>>(nrun, flag.p, rank.p, start2.p);
  Buf<int32_t> dcol(pool, nent);
  Buf<float> dval(pool, nent);
  EVOC_HIP(hipMemsetAsync(cnt.p, 0, (np + 1) * sizeof(int64_t), s));
  Buf<float> accs(pool, nrun);
  cg_gather_kernel<<<(unsigned)ceil_div(nrun, 256), 256, 0, s>>>(nrun, idx2.p, acc.p, accs.p);
  cg_entry_kernel<<<(unsigned)ceil_div(nent, 256), 256, 0, s>>>(
      nent, nrun, np, start2.p, key2.p, accs.p, dcol.p, dval.p, cnt.p);
  exclusive_scan_i64(cnt.p, dptr.p, np + 1, s, pool);
  EVOC_HIP(hipGetLastError());
  if (!keep || host_copy) {
    ccol.resize(nent);
    cval.resize(nent);
    EVOC_HIP(hipMemcpyAsync(cptr.data(), dptr.p, (np + 1) * sizeof(int64_t),
                            hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(ccol.data(), dcol.p, nent * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    EVOC_HIP(hipMemcpyAsync(cval.data(), dval.p, nent * sizeof(float), hipMemcpyDeviceToHost, s));
  }
  EVOC_HIP(hipStreamSynchronize(s));
  if (keep) {
    keep->n = np;
    keep->nnz = nent;
    keep->indptr = std::move(dptr);
    keep->indices = std::move(dcol);
    keep->weights = std::move(dval);
  }
}

void coarse_slots_device(const Graph& g, Graph& out, hipStream_t s, DevPool& pool) {
  const int64_t np = g.n, nnz = g.nnz;
  out.n = np;
  out.indptr.alloc(pool, np + 1);
  Buf<int64_t> deg(pool, np + 1);
  EVOC_HIP(hipMemsetAsync(deg.p, 0, (np + 1) * sizeof(int64_t), s));
  if (nnz == 0 || np == 0) {
    exclusive_scan_i64(deg.p, out.indptr.p, np + 1, s, pool);
    out.nnz = 0;
    out.indices.alloc(pool, 1);
    out.weights.alloc(pool, 1);
    return;
  }
  EVOC_CHECK(2 * nnz < ((int64_t)1 << 32), "coarse embedding: more than 2^31 graph entries");
  Buf<uint64_t> key(pool, 2 * nnz);
  Buf<uint32_t> idx(pool, 2 * nnz);
  Buf<int32_t> erow(pool, nnz);
  slot_key_kernel<<<(unsigned)ceil_div(np, LP_WAVES), 64 * LP_WAVES, 0, s>>>(
      g.indptr.p, g.indices.p, np, key.p, idx.p, erow.p, deg.p);
  radix_sort_pairs(key.p, idx.p, 2 * nnz, bit_length((uint64_t)(2 * np) * (uint64_t)np), s, pool);
  exclusive_scan_i64(deg.p, out.indptr.p, np + 1, s, pool);
  int64_t m = 0;
  EVOC_HIP(hipMemcpyAsync(&m, out.indptr.p + np, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
  out.nnz = m;
  out.indices.alloc(pool, std::max<int64_t>(m, 1));
  out.weights.alloc(pool, std::max<int64_t>(m, 1));
  if (m > 0)
    slot_gather_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, s>>>(m, idx.p, g.indices.p, erow.p,
                                                                 g.weights.p, out.indices.p,
                                                                 out.weights.p);
  EVOC_HIP(hipGetLastError());
}

}  // namespace evoc
