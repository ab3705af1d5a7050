// Total-persistence curve in fp64 (compute_total_persistence,
// src/clustering/persistence.rs:135-173): one thread per size bin sums the
// contributions of the barcodes that cover it, in barcode order, so the value equals
// the sequential fp64 accumulation bit for bit.
#include "kernels.h"

namespace evoc {
namespace {

__global__ __launch_bounds__(256) void bins_kernel(const int64_t* __restrict__ bpos,
                                                   const int64_t* __restrict__ dpos,
                                                   const double* __restrict__ contrib,
                                                   int64_t nbc, int64_t nbins,
                                                   double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbins) return;
  double s = 0.0;
  for (int64_t b = 0; b < nbc; ++b)
    if (bpos[b] <= k && k < dpos[b]) s = s + contrib[b];
  out[k] = s;
}

}  // namespace

void persistence_bins_device(const int64_t* h_bpos, const int64_t* h_dpos,
                             const double* h_contrib, int64_t nbc, int64_t nbins, double* h_out,
                             hipStream_t s, DevPool& pool) {
  if (nbins <= 0) return;
  Buf<int64_t> bp(pool, nbc), dp(pool, nbc);
  Buf<double> cb(pool, nbc), out(pool, nbins);
  if (nbc > 0) {
    EVOC_HIP(hipMemcpyAsync(bp.p, h_bpos, nbc * sizeof(int64_t), hipMemcpyHostToDevice, s));
    EVOC_HIP(hipMemcpyAsync(dp.p, h_dpos, nbc * sizeof(int64_t), hipMemcpyHostToDevice, s));
    EVOC_HIP(hipMemcpyAsync(cb.p, h_contrib, nbc * sizeof(double), hipMemcpyHostToDevice, s));
  }
  bins_kernel<<<(unsigned)ceil_div(nbins, 256), 256, 0, s>>>(bp.p, dp.p, cb.p, nbc, nbins, out.p);
  EVOC_HIP(hipGetLastError());
  EVOC_HIP(hipMemcpyAsync(h_out, out.p, nbins * sizeof(double), hipMemcpyDeviceToHost, s));
  EVOC_HIP(hipStreamSynchronize(s));
}

}  // namespace evoc
