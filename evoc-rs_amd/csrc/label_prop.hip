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
