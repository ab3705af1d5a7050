// Stable LSD radix sort of (uint64 key, uint32 value) pairs, 8 bits per pass.
// Per pass: block digit histograms -> digit-major exclusive scan -> stable scatter
// (wave-level digit matching with ballots, so equal digits keep their input order).
#include "kernels.h"

namespace evoc {
namespace {

constexpr int RS_T = 256;
constexpr int RS_ROUNDS = 8;
constexpr int RS_TILE = RS_T * RS_ROUNDS;

__global__ __launch_bounds__(RS_T) void rs_hist(const uint64_t* __restrict__ key, int64_t n,
                                                int shift, int64_t nblocks,
                                                int64_t* __restrict__ hist) {
  __shared__ unsigned h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; ++r) {
    const int64_t i = base + r * RS_T + threadIdx.x;
    if (i < n) atomicAdd(&h[(key[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_T) void rs_scatter(const uint64_t* __restrict__ key,
                                                   const uint32_t* __restrict__ val, int64_t n,
                                                   int shift, int64_t nblocks,
                                                   const int64_t* __restrict__ off,
                                                   uint64_t* __restrict__ okey,
                                                   uint32_t* __restrict__ oval) {
  __shared__ int64_t boff[256];
  __shared__ unsigned wcnt[4][256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  boff[tid] = off[(int64_t)tid * nblocks + blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < RS_ROUNDS; ++r) {
    for (int w = 0; w < 4; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    const int64_t i = base + r * RS_T + tid;
    const bool have = i < n;
    const uint64_t k = have ? key[i] : 0;
    const unsigned dg = have ? (unsigned)((k >> shift) & 255u) : 256u;
    // lanes of this wave holding the same digit
    unsigned long long peers = __ballot(have);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (dg >> b) & 1u;
      const unsigned long long m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const int rank = __popcll(peers & lt);
    if (have && rank == 0) wcnt[wave][dg] = (unsigned)__popcll(peers);
    __syncthreads();
    if (have) {
      int64_t pos = boff[dg] + rank;
      for (int w = 0; w < wave; ++w) pos += wcnt[w][dg];
      okey[pos] = k;
      oval[pos] = val[i];
    }
    __syncthreads();
    boff[tid] += wcnt[0][tid] + wcnt[1][tid] + wcnt[2][tid] + wcnt[3][tid];
    __syncthreads();
  }
}

}  // namespace

void radix_sort_pairs(uint64_t* d_key, uint32_t* d_val, int64_t n, int key_bits, hipStream_t s,
                      DevPool& pool) {
  if (n <= 1) return;
  const int64_t nb = ceil_div(n, RS_TILE);
  Buf<uint64_t> k2(pool, n);
  Buf<uint32_t> v2(pool, n);
  Buf<int64_t> hist(pool, 256 * nb), off(pool, 256 * nb);
  uint64_t* ka = d_key;
  uint32_t* va = d_val;
  uint64_t* kb = k2.p;
  uint32_t* vb = v2.p;
  const int passes = (key_bits + 7) / 8;
  for (int p = 0; p < passes; ++p) {
    rs_hist<<<(unsigned)nb, RS_T, 0, s>>>(ka, n, 8 * p, nb, hist.p);
    exclusive_scan_i64(hist.p, off.p, 256 * nb, s, pool);
    rs_scatter<<<(unsigned)nb, RS_T, 0, s>>>(ka, va, n, 8 * p, nb, off.p, kb, vb);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  EVOC_HIP(hipGetLastError());
  if (ka != d_key) {
    EVOC_HIP(hipMemcpyAsync(d_key, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
    EVOC_HIP(hipMemcpyAsync(d_val, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
}

}  // namespace evoc
