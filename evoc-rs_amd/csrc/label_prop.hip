// Device parts of the label-propagation initial layout (label_propagation_init,
// src/graph/label_prop.rs:537-564).
//
// partition_means: the data coarsening of one level, (R^T_norm)_l1 X
// (src/utils/sparse.rs:479-501 row L1 normalisation, then the sparse-dense product).
// Row a of the result is coef[a] * sum of the member rows of partition a, accumulated
// member by member in ascending node order, each term rounded as coef * x before the
// add -- the order the sparse product walks a CSR row -- so the f32 values equal the
// host formulation bit for bit. One workgroup per partition, one thread per column
// stripe; eight member rows are loaded ahead of the dependent adds.
#include "kernels.h"

namespace evoc {
namespace {

constexpr int PM_THREADS = 256;
constexpr int PM_AHEAD = 8;

__global__ __launch_bounds__(PM_THREADS) void partition_means_kernel(
    const float* __restrict__ X, int64_t ld, int64_t d, const int64_t* __restrict__ mptr,
    const int32_t* __restrict__ mem, const float* __restrict__ coef, float* __restrict__ out) {
  const int64_t a = blockIdx.x;
  const int64_t m0 = mptr[a], m1 = mptr[a + 1];
  const float cf = coef[a];
  for (int64_t c = threadIdx.x; c < d; c += PM_THREADS) {
    float acc = 0.0f;
    int64_t m = m0;
    for (; m + PM_AHEAD <= m1; m += PM_AHEAD) {
      float v[PM_AHEAD];
#pragma unroll
      for (int u = 0; u < PM_AHEAD; ++u) v[u] = X[(int64_t)mem[m + u] * ld + c];
#pragma unroll
      for (int u = 0; u < PM_AHEAD; ++u) {
        const float t = cf * v[u];
        acc = acc + t;
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
  partition_means_kernel<<<(unsigned)np, PM_THREADS, 0, s>>>(dX, ld, d, mptr.p, mem.p, coef.p,
                                                              out.p);
  EVOC_HIP(hipGetLastError());
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
// running best starting at 1.0, ties to the later label), or stays unlabelled.
__global__ __launch_bounds__(64 * LP_WAVES) void lp_step_kernel(
    const int64_t* __restrict__ indptr, const int32_t* __restrict__ col,
    const float* __restrict__ val, int64_t n, const int64_t* __restrict__ lab_in,
    int64_t* __restrict__ lab_out) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * LP_WAVES + (threadIdx.x >> 6);
  if (i >= n) return;  // whole wavefronts
  const int64_t li0 = lab_in[i];
  if (li0 >= 0) {
    if (lane == 0) lab_out[i] = li0;
    return;
  }
  const int64_t s0 = indptr[i], s1 = indptr[i + 1];
  float top = 1.0f;
  int64_t best = -1;
  for (int64_t pb = s0; pb < s1; pb += 64) {
    const int64_t p = pb + lane;
    const bool pv = p < s1;
    const int64_t lp = pv ? lab_in[col[p]] : -1;
    bool first = lp >= 0;
    float vw = pv ? val[p] : 0.0f;
    for (int64_t qb = s0; qb < s1; qb += 64) {
      const int64_t q = qb + lane;
      const int64_t lq = q < s1 ? lab_in[col[q]] : -1;
      const float wq = q < s1 ? val[q] : 0.0f;
      const int tn = (int)min<int64_t>(64, s1 - qb);
      for (int t = 0; t < tn; ++t) {
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

// centring: per coordinate the mean of res summed in node order, then x0.1
__global__ void exp_colsum_kernel(const float* __restrict__ res, int64_t n, int dim,
                                  float inv_n, float* __restrict__ mean) {
  const int c = threadIdx.x;
  if (c >= dim) return;
  float sum = 0.0f;
  int64_t i = 0;
  for (; i + 8 <= n; i += 8) {  // loads ahead of the dependent adds
    float x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = res[(i + u) * dim + c];
#pragma unroll
    for (int u = 0; u < 8; ++u) sum = sum + x[u];
  }
  for (; i < n; ++i) sum = sum + res[i * dim + c];
  mean[c] = sum * inv_n;
}
__global__ void exp_center_kernel(float* __restrict__ res, int64_t n, int dim,
                                  const float* __restrict__ mean, float* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * dim) return;
  const float x = res[q] - mean[q % dim];
  out[q] = 0.1f * x;
}

}  // namespace

void lp_steps_device(const int64_t* d_indptr, const int32_t* d_col, const float* d_val, int64_t n,
                     int n_iter, int64_t* h_lab, hipStream_t s, DevPool& pool) {
  Buf<int64_t> a(pool, n), b(pool, n);
  EVOC_HIP(hipMemcpyAsync(a.p, h_lab, n * sizeof(int64_t), hipMemcpyHostToDevice, s));
  const unsigned g = (unsigned)ceil_div(n, LP_WAVES);
  for (int it = 0; it < n_iter; ++it) {
    lp_step_kernel<<<g, 64 * LP_WAVES, 0, s>>>(d_indptr, d_col, d_val, n, a.p, b.p);
    std::swap(a.p, b.p);
  }
  EVOC_HIP(hipGetLastError());
  EVOC_HIP(hipMemcpyAsync(h_lab, a.p, n * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
}

void expansion_device(const int64_t* d_indptr, const int32_t* d_col, const float* d_val,
                      int64_t n, int64_t nnz, int dim, const int32_t* h_part, const float* h_sa,
                      int64_t np, const float* h_layout, float* d_out, hipStream_t s,
                      DevPool& pool) {
  const int64_t m = std::max<int64_t>(nnz, 1);
  Buf<int32_t> part(pool, n), ta(pool, m), tf(pool, m), ra(pool, m);
  Buf<float> sa(pool, np), layout(pool, np * dim), tv(pool, m), tacc(pool, m), rv(pool, m),
      res(pool, n * dim), mean(pool, dim);
  EVOC_HIP(hipMemcpyAsync(part.p, h_part, n * sizeof(int32_t), hipMemcpyHostToDevice, s));
  EVOC_HIP(hipMemcpyAsync(sa.p, h_sa, np * sizeof(float), hipMemcpyHostToDevice, s));
  EVOC_HIP(hipMemcpyAsync(layout.p, h_layout, np * dim * sizeof(float), hipMemcpyHostToDevice, s));
  const unsigned g = (unsigned)ceil_div(n, LP_WAVES);
  exp_terms_kernel<<<g, 64 * LP_WAVES, 0, s>>>(d_indptr, d_col, d_val, n, part.p, sa.p, ta.p,
                                               tv.p);
  exp_sums_kernel<<<g, 64 * LP_WAVES, 0, s>>>(d_indptr, n, ta.p, tv.p, tf.p, tacc.p);
  exp_node_kernel<<<g, 64 * LP_WAVES, 0, s>>>(d_indptr, n, dim, part.p, sa.p, layout.p, ta.p, tf.p,
                                              tacc.p, ra.p, rv.p, res.p);
  exp_colsum_kernel<<<1, 64, 0, s>>>(res.p, n, dim, (float)(1.0 / (double)n), mean.p);
  exp_center_kernel<<<(unsigned)ceil_div(n * dim, 256), 256, 0, s>>>(res.p, n, dim, mean.p, d_out);
  EVOC_HIP(hipGetLastError());
  EVOC_HIP(hipStreamSynchronize(s));
}

}  // namespace evoc
