// Device exclusive prefix sum over int64 (three passes: block totals, scan of the
// totals, block-local scan plus offset). Deterministic: integer arithmetic.
#include "kernels.h"

namespace evoc {
namespace {

constexpr int SCAN_T = 256;
constexpr int SCAN_PER = 4;
constexpr int SCAN_CHUNK = SCAN_T * SCAN_PER;

// inclusive scan of one value per thread across a 256-thread block
__device__ int64_t block_inclusive(int64_t v, int64_t* sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int64_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) sh[wave] = v;
  __syncthreads();
  int64_t add = 0;
  for (int w = 0; w < wave; ++w) add += sh[w];
  __syncthreads();
  return v + add;
}

__global__ __launch_bounds__(SCAN_T) void scan_totals(const int64_t* __restrict__ in, int64_t n,
                                                      int64_t* __restrict__ totals) {
  __shared__ int64_t sh[4];
  const int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK;
  int64_t s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_PER; ++j) {
    int64_t i = base + threadIdx.x * SCAN_PER + j;
    if (i < n) s += in[i];
  }
  int64_t inc = block_inclusive(s, sh);
  if (threadIdx.x == SCAN_T - 1) totals[blockIdx.x] = inc;
}

__global__ __launch_bounds__(SCAN_T) void scan_of_totals(int64_t* __restrict__ totals, int64_t nb) {
  __shared__ int64_t sh[4];
  int64_t carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += SCAN_T) {
    int64_t i = b0 + threadIdx.x;
    int64_t v = i < nb ? totals[i] : 0;
    int64_t inc = block_inclusive(v, sh);
    if (i < nb) totals[i] = carry + inc - v;  // exclusive
    __shared__ int64_t last;
    if (threadIdx.x == SCAN_T - 1) last = inc;
    __syncthreads();
    carry += last;
    __syncthreads();
  }
}

__global__ __launch_bounds__(SCAN_T) void scan_apply(const int64_t* __restrict__ in, int64_t n,
                                                     const int64_t* __restrict__ totals,
                                                     int64_t* __restrict__ out) {
  __shared__ int64_t sh[4];
  const int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK;
  int64_t v[SCAN_PER];
  int64_t s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_PER; ++j) {
    int64_t i = base + threadIdx.x * SCAN_PER + j;
    v[j] = i < n ? in[i] : 0;
    s += v[j];
  }
  int64_t excl = block_inclusive(s, sh) - s + totals[blockIdx.x];
#pragma unroll
  for (int j = 0; j < SCAN_PER; ++j) {
    int64_t i = base + threadIdx.x * SCAN_PER + j;
    if (i < n) out[i] = excl;
    excl += v[j];
  }
}

}  // namespace

void exclusive_scan_i64(const int64_t* d_in, int64_t* d_out, int64_t n, hipStream_t s,
                        DevPool& pool) {
  if (n <= 0) return;
  const int64_t nb = ceil_div(n, SCAN_CHUNK);
  Buf<int64_t> totals(pool, nb);
  scan_totals<<<(unsigned)nb, SCAN_T, 0, s>>>(d_in, n, totals.p);
  scan_of_totals<<<1, SCAN_T, 0, s>>>(totals.p, nb);
  scan_apply<<<(unsigned)nb, SCAN_T, 0, s>>>(d_in, n, totals.p, d_out);
  EVOC_HIP(hipGetLastError());
}

}  // namespace evoc
