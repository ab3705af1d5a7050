// Input validation on the device: the pipeline rejects a matrix holding NaN or +-inf
// before any stage reads it (the reference's distance and weight arithmetic is only
// defined on finite values; a non-finite row would otherwise make the MST a forest).
#include "kernels.h"

namespace evoc {
namespace {

// One wavefront per row, rows strided over the grid; a wavefront that saw a non-finite
// value sets the flag with one atomic.
__global__ __launch_bounds__(256) void finite_rows_kernel(const float* __restrict__ X, int64_t n,
                                                          int d, int64_t ld,
                                                          unsigned* __restrict__ flag) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  bool bad = false;
  for (int64_t row = (int64_t)blockIdx.x * 4 + wave; row < n; row += nwaves) {
    const float* r = X + row * ld;
    for (int t = lane; t < d; t += 64) bad |= !__builtin_isfinite(r[t]);
  }
  if (__any(bad) && lane == 0) atomicOr(flag, 1u);
}

__global__ void zero_u32_kernel(unsigned* p) { *p = 0u; }

}  // namespace

void check_finite_device(const float* dX, int64_t n, int d, int64_t ld, const char* what,
                         hipStream_t s, DevPool& pool) {
  if (n <= 0 || d <= 0) return;
  Buf<unsigned> flag(pool, 1);
  zero_u32_kernel<<<1, 1, 0, s>>>(flag.p);
  const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(n, 4), 8192);
  finite_rows_kernel<<<grid, 256, 0, s>>>(dX, n, d, ld, flag.p);
  EVOC_HIP(hipGetLastError());
  unsigned h = 0;
  EVOC_HIP(hipMemcpyAsync(&h, flag.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
  EVOC_CHECK(h == 0, std::string(what) + " contains a non-finite value");
}

}  // namespace evoc
